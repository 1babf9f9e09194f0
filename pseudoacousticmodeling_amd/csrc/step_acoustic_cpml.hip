// step_acoustic_cpml.hip -- k_step instances for acoustic physics with the cpml boundary.
#include "pam_variants.cuh"

namespace pam {
int launch_step_acoustic_cpml(int tile, const StepArgs &a, hipStream_t st) {
    return launch_step_p<0, true>(tile, a, st);
}
}  // namespace pam
