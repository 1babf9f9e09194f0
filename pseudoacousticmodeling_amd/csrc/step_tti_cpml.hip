// step_tti_cpml.hip -- k_step instances for tti physics with the cpml boundary.
#include "pam_variants.cuh"

namespace pam {
int launch_step_tti_cpml(int tile, const StepArgs &a, hipStream_t st) {
    return launch_step_p<2, true>(tile, a, st);
}
}  // namespace pam
