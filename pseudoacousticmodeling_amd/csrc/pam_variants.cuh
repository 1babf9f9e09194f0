// pam_variants.cuh -- the table of step-kernel variants the autotuner and
// PAM_OPT_TILE choose from: 0..2 the LDS-tiled k_tile (16, 32, 64 rows per
// block), 3 the register z-march k_step (4 cells per lane, 16 rows per
// wavefront), 4..6 the LDS-ring z-march k_march (7, 14, 28 chunks of 16 rows
// per block), 7..9 the pipelined 16-row tile k_tile_pipe (interior tiles in
// groups of 2, 4 or 8 shots), 10..11 the same with groups of 4 or 8 whose
// edge tiles update the CPML psi themselves (no pre-pass), 12 the temporal
// blocking k_tblock (tblock_steps(phys) steps per launch, Cerjan, whole forward runs
// only: a single step of it is variant 0's).
#pragma once
#include "pam_step.cuh"
#include "pam_tile.cuh"
#include "pam_march.cuh"

namespace pam {

static constexpr TileShape kTiles[] = {{0, 16}, {0, 32}, {0, 64}, {4, 16}, {-1, 7}, {-1, 14}, {-1, 28}, {-2, 2}, {-2, 4}, {-2, 8}, {-3, 4}, {-3, 8}, {-4, 0}};
static constexpr int kNumTiles = (int)(sizeof(kTiles) / sizeof(kTiles[0]));

template <int PHYS, bool CPML>
static int launch_step_p(int tile, const StepArgs &a, hipStream_t st, const hipEvent_t *ev) {
    switch (tile) {
        case 0: return launch_tile_t<PHYS, CPML, 16, 1>(a, st, ev);
        case 1: return launch_tile_t<PHYS, CPML, 32, 1>(a, st, ev);
        case 2: return launch_tile_t<PHYS, CPML, 64, 1>(a, st, ev);
        case 3: return launch_step_t<PHYS, CPML, 4, 16>(a, st, ev);
        case 4: return launch_march_t<PHYS, CPML, 7>(a, st, ev);
        case 5: return launch_march_t<PHYS, CPML, 14>(a, st, ev);
        case 6: return launch_march_t<PHYS, CPML, 28>(a, st, ev);
        case 7: return launch_tile_pipe_t<PHYS, CPML, 2, false>(a, st, ev);
        case 8: return launch_tile_pipe_t<PHYS, CPML, 4, false>(a, st, ev);
        case 9: return launch_tile_pipe_t<PHYS, CPML, 8, false>(a, st, ev);
        case 10: return launch_tile_pipe_t<PHYS, CPML, 4, true>(a, st, ev);
        case 11: return launch_tile_pipe_t<PHYS, CPML, 8, true>(a, st, ev);
        case 12: return launch_tile_t<PHYS, CPML, 16, 1>(a, st, ev);
        default: return (int)hipErrorInvalidValue;
    }
}

// the group launch exists for the 16-row tile (its psi pass through the table
// pboff) and the z-march variants
template <int PHYS, bool CPML>
static int launch_multi_p(int tile, const StepArgs *list, const int *boff, int ngrid, int nblocks, int k,
                          const int *pboff, int npblocks, hipStream_t st) {
    switch (tile) {
        case 0: return launch_tile_multi_t<PHYS, CPML, 16>(list, boff, ngrid, nblocks, k, pboff, npblocks, st);
        case 4: return launch_march_multi_t<PHYS, CPML, 7>(list, boff, ngrid, nblocks, k, st);
        case 5: return launch_march_multi_t<PHYS, CPML, 14>(list, boff, ngrid, nblocks, k, st);
        case 6: return launch_march_multi_t<PHYS, CPML, 28>(list, boff, ngrid, nblocks, k, st);
        default: return (int)hipErrorInvalidValue;
    }
}

}  // namespace pam
