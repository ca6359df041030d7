// Persistent path-tracker kernel: one path per lane, flattened endgame state machine.
//
// Compiled by hipRTC together with the generated system functions of one homotopy
// (HCG_N, HCG_KIND, sysG_* / sysF_*).  Each lane repeatedly takes a path index from a
// device work queue (atomic counter), tracks it with one tracker step per loop
// iteration, and writes its PathResult record when the path ends; finished lanes of a
// wave take their next paths together (hcg_sched.h).
//
// Algorithm (restated from the reference, same arithmetic contract as the oracle):
//   tracker            src/tracker.jl:517-996     predictor  src/predictor.jl:124-371
//   Newton corrector   src/newton_corrector.jl:55-286
//   linear algebra     src/linear_algebra.jl:88-885   weighted norm  src/norm.jl:101-234
//   valuation          src/valuation.jl:38-228    endgame  src/endgame_tracker.jl:260-911
#pragma once
#include "hcg_arith.h"
#include "hcgpu_types.h"
#include "hcg_kargs.h"
#include "hcg_sched.h"
#include "hcg_mlp_infer.h"
#include "hcg_tracker_body.h"

extern "C" __global__ void __launch_bounds__(64, HCG_WAVES_PER_EU) hcg_track_kernel(hcg_kargs a) {
    __shared__ unsigned short mlp_lds[HCG_MLP_NPARAMS];
    const unsigned short* mw = stage_mlp(a.mlp, a.opts, mlp_lds);
    Lane L;
    lane_setup(L, a, mw);
    LaneTotals T = {0, 0, 0, 0.0};
    const int refill = refill_threshold(a.opts);
    int phase = 0;  // 0: idle, 1: tracking
    int k = -1;
    while (true) {
        const unsigned long long fetch = wave_refill_mask(phase == 0, refill);
        if (phase == 0) {
            if (!fetch) continue;  // wait for more idle lanes of this wave
            k = wave_fetch(fetch, a.next_path);
            if (k >= a.npaths) break;
            if (!lane_begin(L, a, k)) continue;
            phase = 1;
        } else {
            lane_iter(L);
        }
        if (lane_ended(L)) {
            lane_finish(L, a, k, T);
            phase = 0;
        }
    }
    atomicAdd(&a.counters[0], (unsigned long long)T.steps);
    atomicAdd(&a.counters[1], (unsigned long long)T.newton);
    atomicAdd(&a.counters[2], (unsigned long long)T.flops);
    atomicAdd(&a.counters[3], (unsigned long long)T.lacc);
}
