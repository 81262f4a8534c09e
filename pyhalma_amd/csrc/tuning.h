// Process-wide tuning of the FP32 pair kernels (written by halma_tune_set).
#pragma once

namespace halma {

// Defaults: the autotuner's choice on the 2.1M-particle clump workload (first unbinding
// pass 4.34 ms against 4.99 ms for R=2, SUB=8, 16 waves/CU, 4 items/wave).
struct Tuning {
    int tests_per_lane = 2;   // R: test particles held in registers per lane
    int subtile = 4;          // sources summed in plain FP32 before each Kahan add
    int min_chunk = 512;      // smallest source chunk when a source range is split
    int items_per_wave = 8;   // queue items targeted per resident wave
    int waves_per_cu = 32;    // persistent waves per compute unit
    int sym_max = 65536;      // largest halo on the symmetric kernel in the unbinding
};

extern Tuning g_tuning;

// Rigorous relative error bound of the compensated FP32 sum for subtile SUB:
// per term 5.5u (FP32 differences and squares 2.5u after the square root, v_rsq_f32
// <= 1 ulp = 2u, the mass product u), plain sub-tile sums (SUB-1)u, Kahan 2u, plus
// FP64 merge and higher-order terms (covered by the extra 0.5u and 1e-11).
inline double pot_error_bound(int sub) {
    const double u = 5.9604644775390625e-08;  // 2^-24
    return (sub + 7.0) * u + 1e-11;
}

// The symmetric self-potential kernel: test side as above with sub-tiles of 8 terms (a
// Kahan sum per source tile, the tiles' sums added in FP64 and kept as a float pair:
// relative error below 1e-15, inside the 0.5u);
// source side per term 4.5u (differences, squares, v_rsq_f32), the lane's two-term
// product sum 2u, the 6-level reduction tree across the wave 6u, Kahan 2u, + 0.5u:
// 15u either way.
inline double pot_sym_error_bound() { return pot_error_bound(8); }

}  // namespace halma
