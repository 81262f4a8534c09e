// Batched single energy test with exact decisions: the gas test of halo_gas.py:456-476
// (and the star test of halo_properties.py:342-359) over many haloes at once.
//
//   E_i = 1/2 |v_i - v_h|^2 - factor * G * S_i,   bound iff E_i <= 0,
//
// with S_i the pair sum of particle_subroutines.f90:466-514 of test i against its halo's
// sources (any CSR source set: gas cells, dark matter, stars) and v_h the halo's bulk
// velocity given by the caller (halo_gas.py:468 uses the halo's vx, vy, vz).  S comes
// from the persistent FP32 compensated kernel (one launch for every halo); members
// whose |E| lies inside the FP32 error band are re-tested with the FP64 sum in the
// oracle's fixed order (one block each), so the mask equals the FP64 oracle's.
#include <math.h>

#include <algorithm>
#include <vector>

#include "halma_common.h"
#include "potential.h"
#include "staging.h"
#include "sym.h"
#include "tuning.h"

namespace halma {

namespace {

constexpr int kThreads = 256;

struct TestArgs {
    const int* tst_halo;
    const int64_t* src_off;  // device copy, absolute (into sm, sx, sy, sz)
    const float *sm, *sx, *sy, *sz;
    // self form: a halo's sources are its tests (src_off = the test offsets, sm.. the
    // tests' arrays) followed by its extra sources [ext_off[h], ext_off[h+1]) of em..
    const int64_t* ext_off;  // device copy, absolute; null in the plain form
    const float *em, *ex, *ey, *ez;
    const float *tx, *ty, *tz;
    const void *vx, *vy, *vz;
    int vel_t;
    const double* halo_v;
    double cg;        // factor * G
    double bound_rel; // B of the FP32 sums
    const double* S;  // FP32-compensated sums (FP64 values)
    uint8_t* bound;
    int32_t* band;
    unsigned long long* counters;  // [0] band members
    double* s_out;
    int64_t n_tst;
};

__device__ __forceinline__ double kin_of(const TestArgs& a, int64_t i, int h) {
    const double dvx = ldf(a.vx, a.vel_t, i) - a.halo_v[3 * h + 0];
    const double dvy = ldf(a.vy, a.vel_t, i) - a.halo_v[3 * h + 1];
    const double dvz = ldf(a.vz, a.vel_t, i) - a.halo_v[3 * h + 2];
    return 0.5 * ((dvx * dvx + dvy * dvy) + dvz * dvz);
}

__global__ void k_et_tst_halo(const int64_t* __restrict__ tst_off, int H,
                              int* __restrict__ tst_halo) {
    const int h = blockIdx.x;
    if (h >= H) return;
    const int64_t b = tst_off[h] - tst_off[0], e = tst_off[h + 1] - tst_off[0];
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) tst_halo[i] = h;
}

__global__ void k_et_clear(unsigned long long* __restrict__ c) {
    if (threadIdx.x < 4) c[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kThreads) void k_et_decide(TestArgs a) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n_tst) return;
    const int h = a.tst_halo[i];
    const double S = a.S[i];
    const double kin = kin_of(a, i, h);
    const double pot = a.cg * S;
    const double E = kin - pot;
    const double band = a.cg * a.bound_rel * S * 1.01 + 8.0 * 0x1p-53 * (kin + fabs(pot)) +
                        1e-10 * fabs(pot);
    a.bound[i] = E <= 0.0 ? 1 : 0;
    if (a.s_out) a.s_out[i] = S;
    if (fabs(E) <= band) {
        const unsigned long long at = atomicAdd(a.counters, 1ull);
        a.band[at] = static_cast<int32_t>(i);
    }
}

// two source segments read as one list (the tests, then the extra sources)
struct SrcTwo {
    SrcSoA a, b;
    int64_t na;
    __device__ __forceinline__ float4 operator()(int64_t j) const {
        return j < na ? a(j) : b(j - na);
    }
};

__global__ __launch_bounds__(kOrdWidth) void k_et_retest(TestArgs a) {
    __shared__ double sh[kOrdWidth];
    const int64_t nb = static_cast<int64_t>(a.counters[0]);
    for (int64_t w = blockIdx.x; w < nb; w += gridDim.x) {
        const int64_t i = a.band[w];
        const int h = a.tst_halo[i];
        const int64_t s0 = a.src_off[h];
        const int64_t ns = a.src_off[h + 1] - s0;
        SrcSoA src{a.sm + s0, a.sx + s0, a.sy + s0, a.sz + s0};
        double S;
        if (a.ext_off) {
            const int64_t e0 = a.ext_off[h], ne = a.ext_off[h + 1] - e0;
            SrcTwo two{src, SrcSoA{a.em + e0, a.ex + e0, a.ey + e0, a.ez + e0}, ns};
            S = block_pot_f64_ordered(two, ns + ne, a.tx[i], a.ty[i], a.tz[i], sh);
        } else {
            S = block_pot_f64_ordered(src, ns, a.tx[i], a.ty[i], a.tz[i], sh);
        }
        if (threadIdx.x == 0) {
            const double E = kin_of(a, i, h) - a.cg * S;
            a.bound[i] = E <= 0.0 ? 1 : 0;
            if (a.s_out) a.s_out[i] = S;
        }
    }
}

__global__ void k_et_add(double* __restrict__ s, const double* __restrict__ t, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) s[i] = s[i] + t[i];
}

struct EtWs {
    double* S;
    int32_t* band;
    int* tst_halo;
    unsigned long long* counters;
    int64_t* off;  // src_off (H+1) then tst_off (H+1)
    void* pot_ws;
    size_t pot_bytes;
    size_t bytes;
};

EtWs carve_et(void* base, int64_t H, int64_t n_tst, size_t pot_bytes) {
    Carver c(base);
    EtWs w;
    const int64_t t1 = n_tst > 0 ? n_tst : 1;
    w.S = c.take<double>(t1);
    w.band = c.take<int32_t>(t1);
    w.tst_halo = c.take<int>(t1);
    w.counters = c.take<unsigned long long>(4);
    w.off = c.take<int64_t>(2 * (H + 1));
    w.pot_ws = c.take<char>(pot_bytes);
    w.pot_bytes = pot_bytes;
    w.bytes = c.used;
    return w;
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

size_t halma_energy_test_workspace_size(int64_t n_halo, const int64_t* src_off,
                                        const int64_t* tst_off) {
    if (n_halo < 0 || !src_off || !tst_off) return 0;
    const size_t pb = pot_batched_workspace_bytes(n_halo, src_off, tst_off);
    return carve_et(nullptr, n_halo, tst_off[n_halo] - tst_off[0], pb).bytes;
}

int halma_energy_test_batched(int64_t n_halo, const int64_t* src_off, const float* sm,
                              const float* sx, const float* sy, const float* sz,
                              const int64_t* tst_off, const float* tx, const float* ty,
                              const float* tz, const void* vx, const void* vy, const void* vz,
                              int vel_dtype, const double* halo_v, double factor,
                              uint8_t* bound, double* s_out, int64_t* stats, void* ws,
                              size_t ws_bytes, void* stream) {
    if (n_halo < 0 || !src_off || !tst_off)
        return set_error(HALMA_E_ARG, "halma_energy_test_batched: bad halo count or offsets");
    if (vel_dtype != HALMA_F32 && vel_dtype != HALMA_F64)
        return set_error(HALMA_E_ARG, "bad velocity dtype");
    if (!isfinite(factor)) return set_error(HALMA_E_ARG, "factor must be finite");
    if (stats) stats[0] = stats[1] = stats[2] = 0;
    if (n_halo == 0) return 0;
    for (int64_t h = 0; h < n_halo; ++h)
        if (src_off[h + 1] < src_off[h] || tst_off[h + 1] < tst_off[h])
            return set_error(HALMA_E_ARG, "offsets must not decrease");
    const int64_t n_src = src_off[n_halo] - src_off[0];
    const int64_t n_tst = tst_off[n_halo] - tst_off[0];
    if (n_tst == 0) return 0;
    if (!tx || !ty || !tz || !vx || !vy || !vz || !halo_v || !bound ||
        (n_src > 0 && (!sm || !sx || !sy || !sz)))
        return set_error(HALMA_E_ARG, "halma_energy_test_batched: null pointer");
    if (n_tst > INT32_MAX) return set_error(HALMA_E_LIMIT, "too many tests");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t pb = pot_batched_workspace_bytes(n_halo, src_off, tst_off);
    EtWs w = carve_et(ws, n_halo, n_tst, pb);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    std::vector<int64_t> offs(2 * (n_halo + 1));
    std::copy(src_off, src_off + n_halo + 1, offs.begin());
    std::copy(tst_off, tst_off + n_halo + 1, offs.begin() + n_halo + 1);
    int rc = upload_async(w.off, offs.data(), offs.size() * sizeof(int64_t), st);
    if (rc) return rc;
    const int64_t t0 = tst_off[0];
    std::vector<int64_t> to_rel(n_halo + 1);  // tests from zero: S and the mask are dense
    for (int64_t h = 0; h <= n_halo; ++h) to_rel[h] = tst_off[h] - t0;
    // with stats, events bracket the pair kernel (its time is reported)
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (stats)
        for (auto& e : ev) HALMA_HIP(hipEventCreate(&e));
    auto drop_events = [&]() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    rc = pot_batched(n_halo, src_off, sm, sx, sy, sz, to_rel.data(), tx + t0, ty + t0, tz + t0,
                     w.S, nullptr, w.pot_ws, w.pot_bytes, st, ev[0], ev[1]);
    if (rc) {
        drop_events();
        return rc;
    }
    k_et_tst_halo<<<static_cast<unsigned>(n_halo), kThreads, 0, st>>>(
        w.off + n_halo + 1, static_cast<int>(n_halo), w.tst_halo);
    HALMA_LAUNCH_CHECK();
    k_et_clear<<<1, 64, 0, st>>>(w.counters);
    HALMA_LAUNCH_CHECK();
    TestArgs a;
    a.ext_off = nullptr;
    a.em = a.ex = a.ey = a.ez = nullptr;
    a.tst_halo = w.tst_halo;
    a.src_off = w.off;
    a.sm = sm;
    a.sx = sx;
    a.sy = sy;
    a.sz = sz;
    a.tx = tx + t0;
    a.ty = ty + t0;
    a.tz = tz + t0;
    const size_t ve = vel_dtype == HALMA_F64 ? sizeof(double) : sizeof(float);
    a.vx = static_cast<const char*>(vx) + t0 * ve;
    a.vy = static_cast<const char*>(vy) + t0 * ve;
    a.vz = static_cast<const char*>(vz) + t0 * ve;
    a.vel_t = vel_dtype;
    a.halo_v = halo_v;
    double G = 4.3 * 1e-3;
    G *= 1e-6;
    a.cg = factor * G;
    a.bound_rel = pot_error_bound(g_tuning.subtile);
    a.S = w.S;
    a.bound = bound + t0;
    a.band = w.band;
    a.counters = w.counters;
    a.s_out = s_out ? s_out + t0 : nullptr;
    a.n_tst = n_tst;
    k_et_decide<<<static_cast<unsigned>((n_tst + kThreads - 1) / kThreads), kThreads, 0, st>>>(a);
    HALMA_LAUNCH_CHECK();
    const int cus = std::max(device_cu_count(), 1);
    k_et_retest<<<static_cast<unsigned>(cus * 4), kOrdWidth, 0, st>>>(a);
    HALMA_LAUNCH_CHECK();
    if (stats) {
        long long* host = pinned_words(2);
        if (!host) {
            drop_events();
            return set_error(HALMA_E_STATE, "pinned host buffer unavailable");
        }
        hipError_t e = hipMemcpyAsync(host, w.counters, sizeof(long long),
                                      hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
        drop_events();
        if (e != hipSuccess) return hip_fail(e, "energy test statistics", __FILE__, __LINE__);
        stats[0] = host[0];
        stats[1] = static_cast<int64_t>(static_cast<double>(ms) * 1e6);
        int64_t pairs = 0;
        for (int64_t h = 0; h < n_halo; ++h)
            pairs += (src_off[h + 1] - src_off[h]) * (tst_off[h + 1] - tst_off[h]);
        stats[2] = pairs;
    }
    return 0;
}

}  // extern "C"

namespace {

// Workspace of the self form: the plain form's (with the extra sources as its sources),
// then the symmetric self sums' and their FP64 results.
struct EtSelfWs {
    EtWs et;
    double* S_self;   // the extra sources' sums
    int64_t* ext_off; // device copy of the extra-source offsets (absolute)
    void* self_ws;
    size_t self_bytes;
    size_t bytes;
};

EtSelfWs carve_et_self(void* base, int64_t H, const int64_t* tst_off, const int64_t* ext_off) {
    EtSelfWs w;
    const int64_t n_tst = tst_off[H] - tst_off[0];
    const size_t pb = pot_batched_workspace_bytes(H, ext_off, tst_off);
    w.et = carve_et(base, H, n_tst, pb);
    Carver c(base ? static_cast<char*>(base) + w.et.bytes : nullptr);
    w.S_self = c.take<double>(n_tst > 0 ? n_tst : 1);
    w.ext_off = c.take<int64_t>(H + 1);
    w.self_bytes = halma_pot_sum_self_workspace_size(H, tst_off);
    w.self_ws = c.take<char>(w.self_bytes);
    w.bytes = w.et.bytes + c.used;
    return w;
}

}  // namespace

extern "C" {

size_t halma_energy_test_self_workspace_size(int64_t n_halo, const int64_t* tst_off,
                                             const int64_t* ext_off) {
    if (n_halo < 0 || !tst_off || !ext_off) return 0;
    for (int64_t h = 0; h < n_halo; ++h)
        if (tst_off[h + 1] < tst_off[h] || ext_off[h + 1] < ext_off[h]) return 0;
    return carve_et_self(nullptr, n_halo, tst_off, ext_off).bytes;
}

int halma_energy_test_self_batched(int64_t n_halo, const int64_t* tst_off, const float* tm,
                                   const float* tx, const float* ty, const float* tz,
                                   const int64_t* ext_off, const float* em, const float* ex,
                                   const float* ey, const float* ez, const void* vx,
                                   const void* vy, const void* vz, int vel_dtype,
                                   const double* halo_v, double factor, uint8_t* bound,
                                   double* s_out, int64_t* stats, void* ws, size_t ws_bytes,
                                   void* stream) {
    if (n_halo < 0 || !tst_off || !ext_off)
        return set_error(HALMA_E_ARG, "halma_energy_test_self_batched: bad halo count or offsets");
    if (vel_dtype != HALMA_F32 && vel_dtype != HALMA_F64)
        return set_error(HALMA_E_ARG, "bad velocity dtype");
    if (!isfinite(factor)) return set_error(HALMA_E_ARG, "factor must be finite");
    if (stats) stats[0] = stats[1] = stats[2] = 0;
    if (n_halo == 0) return 0;
    for (int64_t h = 0; h < n_halo; ++h)
        if (ext_off[h + 1] < ext_off[h] || tst_off[h + 1] < tst_off[h])
            return set_error(HALMA_E_ARG, "offsets must not decrease");
    const int64_t n_ext = ext_off[n_halo] - ext_off[0];
    const int64_t n_tst = tst_off[n_halo] - tst_off[0];
    if (n_tst == 0) return 0;
    if (!tm || !tx || !ty || !tz || !vx || !vy || !vz || !halo_v || !bound ||
        (n_ext > 0 && (!em || !ex || !ey || !ez)))
        return set_error(HALMA_E_ARG, "halma_energy_test_self_batched: null pointer");
    if (n_tst > INT32_MAX) return set_error(HALMA_E_LIMIT, "too many tests");
    hipStream_t st = static_cast<hipStream_t>(stream);
    EtSelfWs sw = carve_et_self(ws, n_halo, tst_off, ext_off);
    if (ws_bytes < sw.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, sw.bytes);
    const EtWs& w = sw.et;
    // the device offsets: the tests' (the first sources of every halo), then again the
    // tests' (for the test-to-halo map), then the extra sources'
    std::vector<int64_t> offs(2 * (n_halo + 1));
    std::copy(tst_off, tst_off + n_halo + 1, offs.begin());
    std::copy(tst_off, tst_off + n_halo + 1, offs.begin() + n_halo + 1);
    int rc = upload_async(w.off, offs.data(), offs.size() * sizeof(int64_t), st);
    if (rc) return rc;
    const int64_t t0 = tst_off[0];
    std::vector<int64_t> to_rel(n_halo + 1);
    for (int64_t h = 0; h <= n_halo; ++h) to_rel[h] = tst_off[h] - t0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (stats)
        for (auto& e : ev) HALMA_HIP(hipEventCreate(&e));
    auto drop_events = [&]() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    if (ev[0]) HALMA_HIP(hipEventRecord(ev[0], st));
    // the self-term, each pair distance once (symmetric kernel), into S
    rc = halma_pot_sum_self_batched(n_halo, tst_off, tm, tx, ty, tz, w.S - t0, sw.self_ws,
                                    sw.self_bytes, st);
    // the extra sources (dark matter, stars) against the tests, added to S
    if (!rc && n_ext > 0) {
        rc = pot_batched(n_halo, ext_off, em, ex, ey, ez, to_rel.data(), tx + t0, ty + t0,
                         tz + t0, sw.S_self, nullptr, w.pot_ws, w.pot_bytes, st);
        if (!rc) {
            k_et_add<<<static_cast<unsigned>((n_tst + kThreads - 1) / kThreads), kThreads, 0, st>>>(
                w.S, sw.S_self, n_tst);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = hip_fail(e, "k_et_add", __FILE__, __LINE__);
        }
    }
    if (!rc && ev[1]) {
        const hipError_t e = hipEventRecord(ev[1], st);
        if (e != hipSuccess) rc = hip_fail(e, "hipEventRecord", __FILE__, __LINE__);
    }
    if (rc) {
        drop_events();
        return rc;
    }
    // extra-source offsets on the device (absolute), for the re-test
    int64_t* ext_dev = sw.ext_off;
    if (n_ext > 0) {
        rc = upload_async(ext_dev, ext_off, (n_halo + 1) * sizeof(int64_t), st);
        if (rc) {
            drop_events();
            return rc;
        }
    }
    k_et_tst_halo<<<static_cast<unsigned>(n_halo), kThreads, 0, st>>>(
        w.off + n_halo + 1, static_cast<int>(n_halo), w.tst_halo);
    HALMA_LAUNCH_CHECK();
    k_et_clear<<<1, 64, 0, st>>>(w.counters);
    HALMA_LAUNCH_CHECK();
    TestArgs a;
    a.tst_halo = w.tst_halo;
    a.src_off = w.off;
    a.sm = tm;
    a.sx = tx;
    a.sy = ty;
    a.sz = tz;
    a.ext_off = n_ext > 0 ? ext_dev : nullptr;
    a.em = em;
    a.ex = ex;
    a.ey = ey;
    a.ez = ez;
    a.tx = tx + t0;
    a.ty = ty + t0;
    a.tz = tz + t0;
    const size_t ve = vel_dtype == HALMA_F64 ? sizeof(double) : sizeof(float);
    a.vx = static_cast<const char*>(vx) + t0 * ve;
    a.vy = static_cast<const char*>(vy) + t0 * ve;
    a.vz = static_cast<const char*>(vz) + t0 * ve;
    a.vel_t = vel_dtype;
    a.halo_v = halo_v;
    double G = 4.3 * 1e-3;
    G *= 1e-6;
    a.cg = factor * G;
    // both parts of S are within their kernels' relative bounds
    a.bound_rel = std::max(pot_error_bound(g_tuning.subtile), pot_sym_error_bound());
    a.S = w.S;
    a.bound = bound + t0;
    a.band = w.band;
    a.counters = w.counters;
    a.s_out = s_out ? s_out + t0 : nullptr;
    a.n_tst = n_tst;
    k_et_decide<<<static_cast<unsigned>((n_tst + kThreads - 1) / kThreads), kThreads, 0, st>>>(a);
    HALMA_LAUNCH_CHECK();
    const int cus = std::max(device_cu_count(), 1);
    k_et_retest<<<static_cast<unsigned>(cus * 4), kOrdWidth, 0, st>>>(a);
    HALMA_LAUNCH_CHECK();
    if (stats) {
        long long* host = pinned_words(2);
        if (!host) {
            drop_events();
            return set_error(HALMA_E_STATE, "pinned host buffer unavailable");
        }
        hipError_t e = hipMemcpyAsync(host, w.counters, sizeof(long long),
                                      hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
        drop_events();
        if (e != hipSuccess) return hip_fail(e, "energy test statistics", __FILE__, __LINE__);
        stats[0] = host[0];
        stats[1] = static_cast<int64_t>(static_cast<double>(ms) * 1e6);
        int64_t pairs = 0;
        for (int64_t h = 0; h < n_halo; ++h) {
            const int64_t nt = tst_off[h + 1] - tst_off[h];
            pairs += (nt + ext_off[h + 1] - ext_off[h]) * nt;
        }
        stats[2] = pairs;
    }
    return 0;
}

}  // extern "C"
