// step_acoustic_cerjan.hip -- k_step instances for acoustic physics with the cerjan boundary.
#include "pam_variants.cuh"

namespace pam {
int launch_step_acoustic_cerjan(int tile, const StepArgs &a, hipStream_t st, const hipEvent_t *ev) {
    return launch_step_p<0, false>(tile, a, st, ev);
}
int launch_multi_acoustic_cerjan(int tile, const StepArgs *list, const int *boff, int ngrid, int nblocks, int k,
                        const int *pboff, int npblocks, hipStream_t st) {
    return launch_multi_p<0, false>(tile, list, boff, ngrid, nblocks, k, pboff, npblocks, st);
}
}  // namespace pam
