// Pair-potential kernels: sum_j m_j / |x_j - x_i| over the sources whose three
// coordinates all differ from test i's (particle_subroutines.f90:466-514).
//
// FP32 path (the hot one): one persistent launch; every wavefront is an independent
// worker that pops queue items {halo, 64*R-test tile, source chunk} (largest first)
// with one vector atomic.  Each lane holds R test positions in registers; sources are
// wave-uniform and arrive through scalar loads (AoS float4 x,y,z,m), so the inner loop
// is pure VALU: 3 sub, mul + 2 fma, v_rsq_f32, mul, the skip test, add.  SUB terms are
// summed plainly, then Kahan-added into the running (sum, compensation) pair; partials
// of split source ranges are merged in FP64 in chunk order (deterministic).
#include <algorithm>
#include <mutex>
#include <numeric>
#include <utility>
#include <vector>

#include "halma_common.h"
#include "potential.h"
#include "staging.h"
#include "tuning.h"

namespace halma {

// acc + m_j / r_ij for one source s = (x, y, z, m) and one test, with the reference's
// rule that a pair counts only if x_j /= x_i .and. y_j /= y_i .and. z_j /= z_i: a
// difference is zero exactly when the coordinates are equal (FP32 subtraction keeps
// denormals), so the pair counts iff the smallest |difference| is not zero (one
// v_min3_f32 with |.| modifiers; a NaN difference, like the reference's /= on NaN,
// counts: the minimum skips NaN operands and is NaN only when all three are NaN).  The
// mass multiply and the add are one FMA.
template <int R, int SUB>
__global__ __launch_bounds__(256) void k_pot_f32(const float4* __restrict__ src,
                                                 const float4* __restrict__ tst,
                                                 const HaloDesc* __restrict__ halos,
                                                 DynSets dyn,
                                                 const int4* __restrict__ items, int n_items,
                                                 float2* __restrict__ part) {
    const int lane = __lane_id();
    const int n_waves = static_cast<int>(gridDim.x * (blockDim.x / kWave));
    if (dyn.n_items) {
        const int64_t nd = uniform_load(dyn.n_items);
        n_items = static_cast<int>(nd < n_items ? nd : n_items);
    }
    // items round-robin over the waves (the list is in descending item size, so every
    // wave gets a similar share): no queue, no atomics -- a queue head serialises at about
    // 12 ns per claim, which dominated passes of many small items
    const int k0 = __builtin_amdgcn_readfirstlane(
        static_cast<int>(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
    for (int k = k0; k < n_items; k += n_waves) {
        const int4 it = uniform_load(items + k);
        const HaloDesc d = uniform_load(halos + it.x);
        int ns = d.src_len, nt = d.tst_len;
        const float4* sbase = src;
        if (dyn.tst_len) {
            nt = __builtin_amdgcn_readfirstlane(uniform_load(dyn.tst_len + it.x));
            ns = __builtin_amdgcn_readfirstlane(uniform_load(dyn.src_len + it.x));
            if (__builtin_amdgcn_readfirstlane(uniform_load(dyn.src_sel + it.x))) sbase = dyn.src_alt;
        }
        const int chunk =
            dyn.chunk ? __builtin_amdgcn_readfirstlane(uniform_load(dyn.chunk + it.x)) : d.chunk;
        const int i0 = it.y * (R * kWave);
        const int j0 = it.z * chunk;
        const int j1 = min(ns, j0 + chunk);
        if (i0 >= nt || j0 >= j1) continue;

        float xi[R], yi[R], zi[R], as[R], ac[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = min(i0 + r * kWave + lane, nt - 1);
            const float4 p = tst[d.tst_off + i];
            xi[r] = p.x;
            yi[r] = p.y;
            zi[r] = p.z;
            as[r] = 0.0f;
            ac[r] = 0.0f;
        }
        const float4* sp = sbase + d.src_off;
        int j = j0;
        for (; j + SUB <= j1; j += SUB) {
            float ps[R];
#pragma unroll
            for (int r = 0; r < R; ++r) ps[r] = 0.0f;
#pragma unroll
            for (int s = 0; s < SUB; ++s) {
                const float4 q = uniform_load(sp + j + s);
#pragma unroll
                for (int r = 0; r < R; ++r) ps[r] = pair_acc(q, xi[r], yi[r], zi[r], ps[r]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) kahan_add(as[r], ac[r], ps[r]);
        }
        if (j < j1) {
            float ps[R];
#pragma unroll
            for (int r = 0; r < R; ++r) ps[r] = 0.0f;
            for (; j < j1; ++j) {
                const float4 q = uniform_load(sp + j);
#pragma unroll
                for (int r = 0; r < R; ++r) ps[r] = pair_acc(q, xi[r], yi[r], zi[r], ps[r]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) kahan_add(as[r], ac[r], ps[r]);
        }
        float2* out = part + d.part_off + static_cast<int64_t>(it.z) * d.stride;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = i0 + r * kWave + lane;
            if (i < nt) out[i] = make_float2(as[r], ac[r]);
        }
    }
}

__global__ void k_fill_i32(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_pack4(const float* __restrict__ m, const float* __restrict__ x,
                        const float* __restrict__ y, const float* __restrict__ z, int64_t n,
                        float4* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(x[i], y[i], z[i], m ? m[i] : 0.0f);
}

// halo id of every test (tests of halo h are [tst_off, tst_off + tst_len))
__global__ void k_tst_halo(const HaloDesc* __restrict__ halos, int n_halo,
                           int* __restrict__ tst_halo) {
    const int h = blockIdx.x;
    if (h >= n_halo) return;
    const HaloDesc d = halos[h];
    for (int i = threadIdx.x; i < d.tst_len; i += blockDim.x) tst_halo[d.tst_off + i] = h;
}

// Merge per-chunk partials (chunk order) into FP64 sums.
__global__ void k_pot_merge(const HaloDesc* __restrict__ halos, const int* __restrict__ tst_halo,
                            int64_t n_tst, const float2* __restrict__ part,
                            double* __restrict__ out_f64, float* __restrict__ out_f32) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_tst) return;
    const HaloDesc d = halos[tst_halo[g]];
    const int64_t i = g - d.tst_off;
    const int nch = d.src_len > 0 ? (d.src_len + d.chunk - 1) / d.chunk : 0;
    // chunk order, the loads eight at a time in flight
    const float2* pp = part + d.part_off + i;
    double s = 0.0;
    for (int c = 0; c < nch; c += 8) {
        float2 q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            q[u] = c + u < nch ? pp[static_cast<int64_t>(c + u) * d.stride] : make_float2(0.0f, 0.0f);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (c + u < nch) s += static_cast<double>(q[u].x) + static_cast<double>(q[u].y);
    }
    if (out_f64) out_f64[g] = s;
    if (out_f32) out_f32[g] = static_cast<float>(s);
}

// one block per test (grid-stride)
__global__ __launch_bounds__(kOrdWidth) void k_pot_f64_ordered(SrcSoA src, int64_t ns,
                                                               const float* __restrict__ tx,
                                                               const float* __restrict__ ty,
                                                               const float* __restrict__ tz,
                                                               int64_t nt,
                                                               double* __restrict__ out) {
    __shared__ double sh[kOrdWidth];
    for (int64_t i = blockIdx.x; i < nt; i += gridDim.x) {
        const double s = block_pot_f64_ordered(src, ns, tx[i], ty[i], tz[i], sh);
        if (threadIdx.x == 0) out[i] = s;
    }
}

__global__ __launch_bounds__(256) void k_pot_f64_plain(
    const double* __restrict__ m, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, int64_t ns, const double* __restrict__ tx,
    const double* __restrict__ ty, const double* __restrict__ tz, int64_t nt,
    double* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const double xi = tx[i], yi = ty[i], zi = tz[i];
    double acc = 0.0;
    for (int64_t j = 0; j < ns; ++j) {
        if (x[j] != xi && y[j] != yi && z[j] != zi) {
            const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
            acc += m[j] / __builtin_sqrt((dx * dx + dy * dy) + dz * dz);
        }
    }
    out[i] = acc;
}

// ------------------------------------------------------------------ host side ----

std::vector<int32_t> order_desc(const std::vector<int64_t>& key) {
    // one 64-bit sort key per halo: the key's bits above the index bits (descending
    // key = ascending complement), the index below (ties in index order); a plain
    // integer sort of these is several times faster than an indirect stable sort
    const int64_t H = static_cast<int64_t>(key.size());
    int ib = 1;
    while ((int64_t{1} << ib) < H) ++ib;
    int64_t kmax = 0;
    for (int64_t k : key) kmax = std::max(kmax, k);
    int kb = 1;
    while (kb < 63 && (int64_t{1} << kb) <= kmax) ++kb;
    std::vector<int32_t> order(H);
    if (kb + ib <= 64) {
        std::vector<uint64_t> v(H);
        const uint64_t kmask = kb >= 64 ? ~0ull : ((1ull << kb) - 1);
        for (int64_t h = 0; h < H; ++h)
            v[h] = ((~static_cast<uint64_t>(key[h]) & kmask) << ib) | static_cast<uint64_t>(h);
        std::sort(v.begin(), v.end());
        const uint64_t imask = (1ull << ib) - 1;
        for (int64_t q = 0; q < H; ++q) order[q] = static_cast<int32_t>(v[q] & imask);
    } else {
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return key[a] > key[b]; });
    }
    return order;
}

int plan_pot(int64_t H, const int64_t* src_off, const int64_t* tst_off, PotPlan& plan,
             bool with_items, int tests_per_lane) {
    Tuning t = g_tuning;
    if (tests_per_lane > 0) t.tests_per_lane = tests_per_lane;
    plan = PotPlan();
    plan.tests_per_lane = t.tests_per_lane;
    plan.subtile = t.subtile;
    const int tile = t.tests_per_lane * kWave;
    plan.halos.resize(H);
    int cus = device_cu_count();
    if (cus <= 0) cus = 256;  // planning without a device (workspace queries)
    const int64_t waves = static_cast<int64_t>(cus) * t.waves_per_cu;
    plan.grid = static_cast<int>(waves / 4);
    const int64_t target = waves * t.items_per_wave;
    int64_t base_items = 0;
    double work = 0.0;
    for (int64_t h = 0; h < H; ++h) {
        const int64_t ns = src_off[h + 1] - src_off[h];
        const int64_t nt = tst_off[h + 1] - tst_off[h];
        if (ns < 0 || nt < 0 || ns > INT32_MAX || nt > INT32_MAX)
            return set_error(HALMA_E_ARG, "halo %lld: bad offsets", (long long)h);
        base_items += (nt + tile - 1) / tile;
        work += static_cast<double>(ns) * static_cast<double>(nt);
    }
    int64_t part = 0;
    std::vector<int64_t> key(H);
    for (int64_t h = 0; h < H; ++h) {
        const int64_t ns = src_off[h + 1] - src_off[h];
        const int64_t nt = tst_off[h + 1] - tst_off[h];
        const int64_t tiles = (nt + tile - 1) / tile;
        int64_t nch = 1;
        if (base_items < target && tiles > 0 && ns > 0 && work > 0.0) {
            const double want = static_cast<double>(target) *
                                (static_cast<double>(ns) * static_cast<double>(nt) / work) /
                                static_cast<double>(tiles);
            const int64_t max_ch = std::max<int64_t>(1, ns / t.min_chunk);
            nch = std::min<int64_t>(max_ch, std::max<int64_t>(1, static_cast<int64_t>(want)));
        }
        int64_t chunk = ns > 0 ? (ns + nch - 1) / nch : 1;
        chunk = (chunk + t.subtile - 1) / t.subtile * t.subtile;
        nch = ns > 0 ? (ns + chunk - 1) / chunk : 0;
        HaloDesc& d = plan.halos[h];
        d.src_off = src_off[h];
        d.tst_off = tst_off[h];
        d.part_off = part;
        d.src_len = static_cast<int32_t>(ns);
        d.tst_len = static_cast<int32_t>(nt);
        d.chunk = static_cast<int32_t>(chunk);
        d.stride = static_cast<int32_t>(nt);
        part += nch * nt;
        key[h] = std::min<int64_t>(nt, tile) * chunk;
    }
    plan.order = order_desc(key);
    int64_t n_items = 0;
    for (int64_t h = 0; h < H; ++h) {
        const HaloDesc& d = plan.halos[h];
        const int64_t tiles = (d.tst_len + tile - 1) / tile;
        const int64_t nch = d.src_len > 0 ? (d.src_len + d.chunk - 1) / d.chunk : 0;
        n_items += tiles * nch;
    }
    if (n_items > static_cast<int64_t>(INT32_MAX))
        return set_error(HALMA_E_LIMIT, "too many work items");
    plan.n_items = n_items;
    if (with_items) {
        plan.items.reserve(static_cast<size_t>(n_items));
        for (int32_t h : plan.order) {
            const HaloDesc& d = plan.halos[h];
            const int tiles = (d.tst_len + tile - 1) / tile;
            const int nch = d.src_len > 0 ? (d.src_len + d.chunk - 1) / d.chunk : 0;
            for (int c = 0; c < nch; ++c)
                for (int ti = 0; ti < tiles; ++ti)
                    plan.items.push_back(make_int4(static_cast<int>(h), ti, c, 0));
        }
    }
    plan.n_src = src_off[H] - src_off[0];
    plan.n_tst = tst_off[H] - tst_off[0];
    plan.n_part = part;
    return 0;
}

int launch_pack4(const float* m, const float* x, const float* y, const float* z, int64_t n,
                 float4* out, hipStream_t stream) {
    if (n <= 0) return 0;
    const int64_t blocks = (n + 255) / 256;
    k_pack4<<<dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream>>>(m, x, y, z, n, out);
    HALMA_LAUNCH_CHECK();
    return 0;
}

template <int R, int SUB>
static int launch_pot_t(const PotPlan& plan, const float4* src4, const float4* tst4,
                        const HaloDesc* d_halos, DynSets dyn, const int4* d_items,
                        float2* part, hipStream_t stream) {
    const int n_items = static_cast<int>(plan.n_items);
    if (n_items == 0) return 0;
    const int grid = std::max(1, std::min(plan.grid, (n_items + 3) / 4));
    k_pot_f32<R, SUB><<<grid, 256, 0, stream>>>(src4, tst4, d_halos, dyn, d_items, n_items,
                                                part);
    HALMA_LAUNCH_CHECK();
    return 0;
}

int launch_pot_f32(const PotPlan& plan, const float4* src4, const float4* tst4,
                   const HaloDesc* d_halos, DynSets dyn, const int4* d_items, float2* part,
                   hipStream_t stream) {
#define HALMA_POT_CASE(R_, S_)                                                            \
    if (plan.tests_per_lane == R_ && plan.subtile == S_)                                  \
        return launch_pot_t<R_, S_>(plan, src4, tst4, d_halos, dyn, d_items, part, stream);
    HALMA_POT_CASE(1, 4)
    HALMA_POT_CASE(1, 8)
    HALMA_POT_CASE(2, 4)
    HALMA_POT_CASE(2, 8)
    HALMA_POT_CASE(4, 4)
    HALMA_POT_CASE(4, 8)
#undef HALMA_POT_CASE
    return set_error(HALMA_E_STATE, "no kernel for R=%d SUB=%d", plan.tests_per_lane,
                     plan.subtile);
}

// Workspace layout of a batched pair-sum call.
struct PotWs {
    float4* src4;
    float4* tst4;
    HaloDesc* halos;
    int4* items;
    float2* part;
    int* tst_halo;
    size_t bytes;
};

static PotWs carve_pot(void* base, const PotPlan& p) {
    Carver c(base);
    PotWs w;
    w.src4 = c.take<float4>(std::max<int64_t>(p.n_src, 1));
    w.tst4 = c.take<float4>(std::max<int64_t>(p.n_tst, 1));
    w.halos = c.take<HaloDesc>(std::max<size_t>(p.halos.size(), 1));
    w.items = c.take<int4>(std::max<int64_t>(p.n_items, 1));
    w.part = c.take<float2>(std::max<int64_t>(p.n_part, 1));
    w.tst_halo = c.take<int>(std::max<int64_t>(p.n_tst, 1));
    w.bytes = c.used;
    return w;
}

size_t pot_batched_workspace_bytes(int64_t H, const int64_t* src_off, const int64_t* tst_off) {
    PotPlan plan;
    if (plan_pot(H, src_off, tst_off, plan, false)) return 0;
    return carve_pot(nullptr, plan).bytes;
}

int pot_batched(int64_t H, const int64_t* src_off, const float* sm, const float* sx,
                       const float* sy, const float* sz, const int64_t* tst_off,
                       const float* tx, const float* ty, const float* tz, double* out_f64,
                       float* out_f32, void* ws, size_t ws_bytes, hipStream_t stream,
                       hipEvent_t ev_begin, hipEvent_t ev_end) {
    PotPlan plan;
    int rc = plan_pot(H, src_off, tst_off, plan);
    if (rc) return rc;
    PotWs w = carve_pot(ws, plan);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    if (plan.n_tst == 0) return 0;
    const int64_t s0 = src_off[0], t0 = tst_off[0];
    // the descriptors index the packed copies from zero
    for (auto& d : plan.halos) {
        d.src_off -= s0;
        d.tst_off -= t0;
    }
    rc = launch_pack4(sm + s0, sx + s0, sy + s0, sz + s0, plan.n_src, w.src4, stream);
    if (rc) return rc;
    rc = launch_pack4(nullptr, tx + t0, ty + t0, tz + t0, plan.n_tst, w.tst4, stream);
    if (rc) return rc;
    rc = upload_async(w.halos, plan.halos.data(), plan.halos.size() * sizeof(HaloDesc), stream);
    if (rc) return rc;
    rc = upload_async(w.items, plan.items.data(), plan.items.size() * sizeof(int4), stream);
    if (rc) return rc;
    k_tst_halo<<<static_cast<unsigned>(H), 256, 0, stream>>>(w.halos, static_cast<int>(H),
                                                             w.tst_halo);
    HALMA_LAUNCH_CHECK();
    if (ev_begin) HALMA_HIP(hipEventRecord(ev_begin, stream));
    rc = launch_pot_f32(plan, w.src4, w.tst4, w.halos, DynSets(), w.items, w.part, stream);
    if (rc) return rc;
    if (ev_end) HALMA_HIP(hipEventRecord(ev_end, stream));
    const int64_t blocks = (plan.n_tst + 255) / 256;
    k_pot_merge<<<static_cast<unsigned>(blocks), 256, 0, stream>>>(
        w.halos, w.tst_halo, plan.n_tst, w.part, out_f64 ? out_f64 + t0 : nullptr,
        out_f32 ? out_f32 + t0 : nullptr);
    HALMA_LAUNCH_CHECK();
    return 0;
}

// ---- the entry point that allocates nothing -----------------------------------------
// Its workspace is this file-scope device array (one copy per GPU, held until the
// process ends).  Host code never uses the array's name as an address: the device
// address comes from hipGetSymbolAddress, once per device, with that device current.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

}  // namespace halma

static __device__ __align__(256) unsigned char scratch[halma::kNoAllocScratchBytes];

namespace halma {

namespace {

// Device address of `scratch` on the current device (looked up on the first call there).
int scratch_address(void** out) {
    static std::mutex mu;
    static std::vector<std::pair<int, void*>> addr;
    int dev = 0;
    HALMA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    for (const auto& e : addr)
        if (e.first == dev) {
            *out = e.second;
            return 0;
        }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    if (!p) return set_error(HALMA_E_STATE, "no device address for the static scratch");
    addr.emplace_back(dev, p);
    *out = p;
    return 0;
}

// The single halo's descriptor and its work items, written on the device (the order of
// plan_pot: chunks outer, test tiles inner), so that nothing is uploaded from the host.
__global__ void k_single_plan(HaloDesc d, int tiles, int n_items, HaloDesc* __restrict__ halos,
                              int4* __restrict__ items) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) halos[0] = d;
    if (k < n_items) items[k] = make_int4(0, k % tiles, k / tiles, 0);
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

double halma_pot_error_bound(void) { return pot_error_bound(g_tuning.subtile); }

size_t halma_pot_sum_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

int halma_pot_sum_f32_noalloc(const float* m, const float* x, const float* y, const float* z,
                              int64_t n_src, const float* tx, const float* ty, const float* tz,
                              int64_t n_test, float* out, double* out_f64, void* stream) {
    if (n_src < 0 || n_test < 0) return set_error(HALMA_E_ARG, "negative size");
    if (n_test > 0 && !out && !out_f64)
        return set_error(HALMA_E_ARG, "halma_pot_sum_f32_noalloc: no output");
    if ((n_src > 0 && (!m || !x || !y || !z)) || (n_test > 0 && (!tx || !ty || !tz)))
        return set_error(HALMA_E_ARG, "halma_pot_sum_f32_noalloc: null data pointer");
    if (n_src > INT32_MAX || n_test > INT32_MAX)
        return set_error(HALMA_E_LIMIT, "at most 2^31-1 sources and tests per halo");
    const int64_t so[2] = {0, n_src}, to[2] = {0, n_test};
    PotPlan plan;
    int rc = plan_pot(1, so, to, plan, false);
    if (rc) return rc;
    const size_t need = carve_pot(nullptr, plan).bytes;
    if (need > kNoAllocScratchBytes)
        return set_error(HALMA_E_LIMIT,
                         "halma_pot_sum_f32_noalloc: %lld sources x %lld tests need %zu bytes of "
                         "scratch, the static scratch holds %zu",
                         (long long)n_src, (long long)n_test, need, kNoAllocScratchBytes);
    if (n_test == 0) return 0;
    void* base = nullptr;
    rc = scratch_address(&base);
    if (rc) return rc;
    PotWs w = carve_pot(base, plan);
    hipStream_t st = static_cast<hipStream_t>(stream);
    rc = launch_pack4(m, x, y, z, plan.n_src, w.src4, st);
    if (rc) return rc;
    rc = launch_pack4(nullptr, tx, ty, tz, plan.n_tst, w.tst4, st);
    if (rc) return rc;
    const int tile = plan.tests_per_lane * kWave;
    const int tiles = static_cast<int>((n_test + tile - 1) / tile);
    const int n_items = static_cast<int>(plan.n_items);
    k_single_plan<<<static_cast<unsigned>(std::max(1, (n_items + 255) / 256)), 256, 0, st>>>(
        plan.halos[0], std::max(tiles, 1), n_items, w.halos, w.items);
    HALMA_LAUNCH_CHECK();
    k_tst_halo<<<1, 256, 0, st>>>(w.halos, 1, w.tst_halo);
    HALMA_LAUNCH_CHECK();
    rc = launch_pot_f32(plan, w.src4, w.tst4, w.halos, DynSets(), w.items, w.part, st);
    if (rc) return rc;
    k_pot_merge<<<static_cast<unsigned>((plan.n_tst + 255) / 256), 256, 0, st>>>(
        w.halos, w.tst_halo, plan.n_tst, w.part, out_f64, out);
    HALMA_LAUNCH_CHECK();
    return 0;
}

size_t halma_pot_sum_batched_workspace_size(int64_t H, const int64_t* src_off,
                                            const int64_t* tst_off) {
    if (H < 0 || !src_off || !tst_off) return 0;
    return pot_batched_workspace_bytes(H, src_off, tst_off);
}

size_t halma_pot_sum_workspace_size(int64_t n_src, int64_t n_test) {
    const int64_t so[2] = {0, n_src}, to[2] = {0, n_test};
    return halma_pot_sum_batched_workspace_size(1, so, to);
}

int halma_pot_sum_batched_f32(int64_t H, const int64_t* src_off, const float* sm,
                              const float* sx, const float* sy, const float* sz,
                              const int64_t* tst_off, const float* tx, const float* ty,
                              const float* tz, double* out_f64, void* ws, size_t ws_bytes,
                              void* stream) {
    if (H < 0 || !src_off || !tst_off || !out_f64)
        return set_error(HALMA_E_ARG, "halma_pot_sum_batched_f32: null argument");
    if (H == 0) return 0;
    for (int64_t h = 0; h < H; ++h)
        if (src_off[h + 1] < src_off[h] || tst_off[h + 1] < tst_off[h])
            return set_error(HALMA_E_ARG, "offsets must be non-decreasing");
    if ((src_off[H] > src_off[0] && (!sm || !sx || !sy || !sz)) ||
        (tst_off[H] > tst_off[0] && (!tx || !ty || !tz)))
        return set_error(HALMA_E_ARG, "halma_pot_sum_batched_f32: null data pointer");
    return pot_batched(H, src_off, sm, sx, sy, sz, tst_off, tx, ty, tz, out_f64, nullptr, ws,
                       ws_bytes, static_cast<hipStream_t>(stream));
}

int halma_pot_sum_f32(const float* m, const float* x, const float* y, const float* z,
                      int64_t n_src, const float* tx, const float* ty, const float* tz,
                      int64_t n_test, float* out, double* out_f64, void* ws, size_t ws_bytes,
                      void* stream) {
    if (n_src < 0 || n_test < 0) return set_error(HALMA_E_ARG, "negative size");
    if (n_test > 0 && !out && !out_f64)
        return set_error(HALMA_E_ARG, "halma_pot_sum_f32: no output");
    if ((n_src > 0 && (!m || !x || !y || !z)) || (n_test > 0 && (!tx || !ty || !tz)))
        return set_error(HALMA_E_ARG, "halma_pot_sum_f32: null data pointer");
    if (n_src > INT32_MAX || n_test > INT32_MAX)
        return set_error(HALMA_E_LIMIT, "at most 2^31-1 sources and tests per halo");
    const int64_t so[2] = {0, n_src}, to[2] = {0, n_test};
    return pot_batched(1, so, m, x, y, z, to, tx, ty, tz, out_f64, out, ws, ws_bytes,
                       static_cast<hipStream_t>(stream));
}

int halma_pot_sum_f64_ordered(const float* m, const float* x, const float* y, const float* z,
                              int64_t n_src, const float* tx, const float* ty,
                              const float* tz, int64_t n_test, double* out, void* stream) {
    if (n_src < 0 || n_test < 0) return set_error(HALMA_E_ARG, "negative size");
    if (n_test == 0) return 0;
    if (!out || !tx || !ty || !tz || (n_src > 0 && (!m || !x || !y || !z)))
        return set_error(HALMA_E_ARG, "halma_pot_sum_f64_ordered: null pointer");
    const int64_t blocks = std::min<int64_t>(n_test, 65536);
    SrcSoA src{m, x, y, z};
    k_pot_f64_ordered<<<static_cast<unsigned>(blocks), kOrdWidth, 0,
                        static_cast<hipStream_t>(stream)>>>(src, n_src, tx, ty, tz, n_test,
                                                             out);
    HALMA_LAUNCH_CHECK();
    return 0;
}

int halma_pot_sum_f64_plain(const double* m, const double* x, const double* y, const double* z,
                            int64_t n_src, const double* tx, const double* ty,
                            const double* tz, int64_t n_test, double* out, void* stream) {
    if (n_src < 0 || n_test < 0) return set_error(HALMA_E_ARG, "negative size");
    if (n_test == 0) return 0;
    if (!out || !tx || !ty || !tz || (n_src > 0 && (!m || !x || !y || !z)))
        return set_error(HALMA_E_ARG, "halma_pot_sum_f64_plain: null pointer");
    const int64_t blocks = (n_test + 255) / 256;
    k_pot_f64_plain<<<static_cast<unsigned>(blocks), 256, 0, static_cast<hipStream_t>(stream)>>>(
        m, x, y, z, n_src, tx, ty, tz, n_test, out);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
