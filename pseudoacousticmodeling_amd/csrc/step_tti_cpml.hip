// step_tti_cpml.hip -- k_step instances for tti physics with the cpml boundary.
#include "pam_variants.cuh"

namespace pam {
int launch_step_tti_cpml(int tile, const StepArgs &a, hipStream_t st, const hipEvent_t *ev) {
    return launch_step_p<2, true>(tile, a, st, ev);
}
int launch_multi_tti_cpml(int tile, const StepArgs *list, const int *boff, int ngrid, int nblocks, int k,
                        const int *pboff, int npblocks, hipStream_t st) {
    return launch_multi_p<2, true>(tile, list, boff, ngrid, nblocks, k, pboff, npblocks, st);
}
}  // namespace pam
