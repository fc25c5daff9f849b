// Stage engine kernels and host logic (see gnt_engine.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gnt_carry.hpp"
#include "gnt_engine.hpp"
#include "gnt_pass.hpp"

namespace gnt {

namespace {

constexpr int TBLOCK = 1024;  // transform kernels: 64 KiB tile, 8 waves per SIMD at two tiles per CU
using X = DevExec<TBLOCK>;
constexpr int SCAN_BLOCK = 1024;

size_t ceil_log2(size_t x) {
    size_t l = 0;
    while (((size_t)1 << l) < x) l++;
    return l;
}

// Largest i with items[i].tile0 <= t (items sorted by tile0, items[0].tile0 == 0).
template <class T>
__device__ __forceinline__ int find_item(const T *__restrict__ items, int n, uint64_t t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (items[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A descriptor every lane of the wave holds the same copy of: read word by word
// through readfirstlane so the compiler keeps it in scalar registers (the
// vector registers are all needed by the tile body).
template <class T>
__device__ __forceinline__ T wave_uniform(const T &v) {
    static_assert(sizeof(T) % 4 == 0, "descriptor of 32-bit words");
    T r;
    const int *src = reinterpret_cast<const int *>(&v);
    int *dst = reinterpret_cast<int *>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return r;
}

template <class G, int KIND>
__global__ __launch_bounds__(TBLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_tpass(
    const TItem *__restrict__ items, int n_items, uint64_t total, uint64_t base, Tw tw) {
    // One tile per workgroup, as the single-product kernels run: a loop over
    // tiles inside the workgroup lets the compiler keep per-thread address
    // arithmetic alive across iterations, which spills the tile bodies.  A
    // launch covers tiles [base, base + gridDim.x) of the group.
    extern __shared__ uint64_t tile[];
    const uint64_t t = base + blockIdx.x;
    if (t >= total) return;
    const TItem it = wave_uniform(items[find_item(items, n_items, t)]);
    if constexpr (KIND == K_FUSED_AB) {
        using H = typename HalfGeom<G>::type;  // half-width tiles of the last-pass shape
        if constexpr (HalfGeom<G>::ok)
            body_fused_ab<X, H>(tile, make_geom<H>(it.d, t - it.tile0), it.x, it.other, it.log_n, tw,
                                StoreTw{it.T, it.log_t});
        return;
    }
    const G g = make_geom<G>(it.d, t - it.tile0);
    const StoreTw st{it.T, it.log_t};
    const int scale_e = it.d.log_m == it.log_n ? inv_scale_e(it.log_n) : 0;  // pass 0 carries 1/N
    if constexpr (KIND == K_FWD) body_fwd<X, G>(tile, g, it.x, it.limbs, it.n_limbs_in, tw, it.T);
    else if constexpr (KIND == K_INV) body_inv<X, G, false, true, false>(tile, g, it.x, tw, nullptr, nullptr, 0, 0, scale_e, st);
    else if constexpr (KIND == K_INV_PLAIN)
        body_inv<X, G, false, false, false>(tile, g, it.x, tw, nullptr, nullptr, 0, 0, scale_e, st);
    else if constexpr (KIND == K_INV_TSTORE)
        body_inv<X, G, false, false, true>(tile, g, it.x, tw, nullptr, nullptr, 0, 0, scale_e, st);
    else if constexpr (KIND == K_INV_SUMS)
        body_inv<X, G, true, true, false>(tile, g, it.x, tw, it.Ls, it.Hs, it.n_coeffs, it.n_limbs, scale_e, st);
    else if constexpr (KIND == K_INV_SUMS_PLAIN)
        body_inv<X, G, true, false, false>(tile, g, it.x, tw, it.Ls, it.Hs, it.n_coeffs, it.n_limbs, scale_e, st);
    else body_fused<X, G>(tile, g, it.x, it.limbs, it.n_limbs_in, it.xa, it.other, it.log_n, tw, it.Ls, it.Hs,
                          it.n_coeffs, it.n_limbs, st);
}


__global__ __launch_bounds__(CARRY_BLOCK) void k_ctiles(const CItem *__restrict__ items, int n_items, uint64_t total) {
    __shared__ uint32_t sh[2 * CARRY_WAVES];
    for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
        const CItem it = wave_uniform(items[find_item(items, n_items, t)]);
        const uint64_t k = t - it.tile0;
        uint32_t gp, hi;
        if (it.mode == C_SUMS) carry_tile(SumsSrc{it.Ls, it.Hs}, it.n_limbs, k, it.out, sh, gp, hi);
        else carry_tile(LinearSrc{it.Ls, it.Hs, it.na, it.nb, it.mode == C_SUB}, it.n_limbs, k, it.out, sh, gp, hi);
        if (threadIdx.x == 0) {
            it.hi[k] = hi;
            it.gp[k] = (unsigned char)gp;
        }
    }
}

// One workgroup per product: tile carry-ins and the trimmed length.
__global__ __launch_bounds__(SCAN_BLOCK) void k_cscan(const CItem *__restrict__ items) {
    __shared__ uint32_t scan[SCAN_BLOCK];
    __shared__ uint32_t maxhi;
    const CItem it = items[blockIdx.x];
    const size_t nt = it.ntiles;
    const size_t per = (nt + SCAN_BLOCK - 1) / SCAN_BLOCK;
    const size_t t0 = (size_t)threadIdx.x * per;
    uint32_t agg = 2, mh = 0;
    for (size_t t = t0; t < t0 + per && t < nt; t++) agg = gp_combine(agg, it.gp[t]);
    scan[threadIdx.x] = agg;
    if (threadIdx.x == 0) maxhi = 0;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        uint32_t v = scan[threadIdx.x];
        uint32_t w = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 2u;
        __syncthreads();
        scan[threadIdx.x] = gp_combine(w, v);
        __syncthreads();
    }
    uint32_t carry = threadIdx.x > 0 ? (scan[threadIdx.x - 1] & 1) : 0;
    for (size_t t = t0; t < t0 + per && t < nt; t++) {
        it.cin[t] = (unsigned char)carry;
        uint32_t gp = it.gp[t];
        // a tile of all-ones limbs that receives a carry becomes zeros (only
        // at the top of a difference, where the carry leaves the result)
        if (!(carry && (gp >> 1))) mh = max(mh, it.hi[t]);
        carry = (gp & 1) | ((gp >> 1) & carry);
    }
    if (mh) atomicMax(&maxhi, mh);
    __syncthreads();
    if (threadIdx.x == 0) *it.out_len = maxhi;
}

// Tiles whose carry-in is 1 add it (one wave per tile).
__global__ __launch_bounds__(CARRY_RIPPLE_BLOCK) void k_cripple(const CItem *__restrict__ items, int n_items,
                                                                uint64_t total) {
    for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
        const CItem it = wave_uniform(items[find_item(items, n_items, t)]);
        const uint64_t k = t - it.tile0;
        if (it.cin[k] == 0) continue;  // uniform per workgroup
        const uint32_t hi = carry_ripple_tile(it.out, it.n_limbs, k);
        if (threadIdx.x == 0 && hi) atomicMax(it.out_len, hi);
    }
}

// After the ripple: each finished item's trimmed length to its pinned host
// word, one thread per item (replaces a 4-byte copy per product).
__global__ __launch_bounds__(256) void k_lens(const CItem *__restrict__ items, int n_items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_items && items[i].len_host) *items[i].len_host = *items[i].out_len;
}

}  // namespace

size_t Prod::cs_bytes() const { return (2 * nt + 15) / 16 * 16 + 4 * nt; }

size_t Prod::scratch_bytes() const {
    if (linear()) return cs_bytes();
    return N * 8 * (sqr() ? 1 : 2) + 2 * nl * 8 + cs_bytes();
}

bool prod_init(Prod &p) {
    p.stage = 0;
    if (p.linear()) {
        p.nl = p.op == OP_ADD ? std::max(p.na, p.nb) + 1 : p.na;
        p.N = p.n_out = 0;
        p.nt = (p.nl + CARRY_TILE - 1) / CARRY_TILE;
        p.nstages = 3;
        return p.nl <= 0xFFFFFFFFull;
    }
    if (p.sqr()) p.nb = p.na;
    p.nl = p.na + p.nb;
    p.n_out = 4 * p.nl - 1;
    int log_n = (int)ceil_log2(p.n_out);
    if (log_n > MAX_LOG_N) return false;
    p.pl = make_plan(log_n);
    p.N = (size_t)1 << log_n;
    p.nt = (p.nl + CARRY_TILE - 1) / CARRY_TILE;
    const int P = p.pl.npass;
    // a multiply with two or more passes runs both last forward passes in the
    // fused stage, as a square does its one
    p.nstages = (p.sqr() || P > 1) ? 2 * P + 2 : 2 * P + 3;
    p.stage = 0;
    return true;
}

// The table the store of inverse pass `pass` (or of the fused stage) applies:
// that of inverse pass pass-1, which runs next.
static void inv_store_table(const Prod &p, int pass, TItem &t) {
    if (pass >= 1 && p.tabs.inv[pass - 1]) {
        t.T = p.tabs.inv[pass - 1];
        t.log_t = p.pl.pass[pass - 1].log_m;
    }
}

static TItem titem(const Prod &p, int pass, uint64_t *x, const uint64_t *limbs, size_t n_in) {
    TItem t{};
    t.x = x;
    t.limbs = limbs;
    t.n_limbs_in = n_in;
    t.n_coeffs = p.n_out;
    t.n_limbs = p.nl;
    t.d = p.pl.pass[pass];
    t.log_n = p.pl.log_n;
    return t;
}

// Forward pass `pass` of operand A or B (packed from its limbs at pass 0).
static void add_fwd(Iteration &it, const Prod &p, int pass, bool b) {
    TItem t = b ? titem(p, pass, p.xb, pass == 0 ? p.db : nullptr, p.nb)
                : titem(p, pass, p.xa, pass == 0 ? p.da : nullptr, p.na);
    t.pad = K_FWD;
    t.T = p.tabs.fwd[pass];
    it.titems.push_back(t);
}

static void add_one_stage(Iteration &it, Prod &p) {
    const int P = p.pl.npass;
    const int s = p.stage++;
    const bool sq = p.sqr();
    uint64_t *Ls = p.lh, *Hs = p.lh + p.nl;
    // forward passes before the fused one
    const bool ab = !sq && P > 1;      // fused stage takes both last forward passes
    const int nfwd = (sq || ab) ? P - 1 : P;  // stages that hold forward passes
    if (p.linear()) {
        CItem c{};
        c.Ls = p.da;
        c.Hs = p.db;
        c.na = p.na;
        c.nb = p.nb;
        c.mode = p.op == OP_ADD ? C_ADD : C_SUB;
        c.out = p.dout;
        c.gp = p.cs;
        c.cin = p.cs + p.nt;
        c.hi = (uint32_t *)(p.cs + (2 * p.nt + 15) / 16 * 16);
        c.out_len = p.dlen;
        c.n_limbs = p.nl;
        c.ntiles = p.nt;
        c.tile0 = (uint64_t)s;  // the carry step, resolved in finalize_iteration
        if (s == 2 && p.len_host) {
            c.len_host = p.len_host;
            it.want_lens = true;
        }
        it.citems.push_back(c);
        return;
    }
    if (s < nfwd) {
        add_fwd(it, p, s, false);
        if (!sq && s < P - 1) add_fwd(it, p, s, true);
        return;
    }
    if (s == nfwd && ab) {
        TItem f = titem(p, P - 1, p.xa, nullptr, 0);
        f.other = p.xb;
        f.pad = K_FUSED_AB;
        inv_store_table(p, P - 1, f);
        it.titems.push_back(f);
        return;
    }
    if (s == nfwd) {
        TItem f = titem(p, P - 1, sq ? p.xa : p.xb, P == 1 ? (sq ? p.da : p.db) : nullptr, sq ? p.na : p.nb);
        f.xa = p.xa;
        f.other = sq ? nullptr : p.xa;
        if (P == 1) { f.Ls = Ls; f.Hs = Hs; }
        f.pad = K_FUSED;
        inv_store_table(p, P - 1, f);
        it.titems.push_back(f);
        return;
    }
    const int inv_stage = s - nfwd - 1;  // 0 .. P-2
    if (inv_stage < P - 1) {
        const int pass = P - 2 - inv_stage;
        TItem v = titem(p, pass, p.xa, nullptr, 0);
        if (pass == 0) { v.Ls = Ls; v.Hs = Hs; }
        if (pass > 0) inv_store_table(p, pass, v);
        // a pass with a table had its twiddles applied by the stage before
        const bool tab = p.tabs.inv[pass] != nullptr;
        if (pass == 0) v.pad = tab ? K_INV_SUMS_PLAIN : K_INV_SUMS;
        else v.pad = !tab ? K_INV : v.T ? K_INV_TSTORE : K_INV_PLAIN;
        it.titems.push_back(v);
        return;
    }
    const int cstep = inv_stage - (P - 1);  // 0, 1, 2
    CItem c{};
    c.Ls = Ls;
    c.Hs = Hs;
    c.out = p.dout;
    c.gp = p.cs;
    c.cin = p.cs + p.nt;
    c.hi = (uint32_t *)(p.cs + (2 * p.nt + 15) / 16 * 16);
    c.out_len = p.dlen;
    c.n_limbs = p.nl;
    c.ntiles = p.nt;
    c.tile0 = (uint64_t)cstep;  // the step, resolved in finalize_iteration
    if (cstep == 2 && p.len_host) {
        c.len_host = p.len_host;
        it.want_lens = true;
    }
    it.citems.push_back(c);
}

// The last transform stage (or a sum's first carry step) takes the three
// carry steps with it: an iteration launches its transform groups before its
// carry groups, and those in step order, on one stream.
int next_stage_span(const Prod &p) {
    const int carry0 = p.nstages - 3;  // first carry step
    return (p.stage == carry0 - 1 || p.stage == carry0) ? p.nstages - p.stage : 1;
}

// A b_late multiply (two or more passes): forward stages of A while B is
// missing, B's forward passes once it is there (one per iteration, beside
// A's), then the usual stages; B's last pass and the fused stage may share an
// iteration (forward groups launch first).
static int fwd_stages(const Prod &p) { return p.pl.npass - 1; }

bool waits_for_b(const Prod &p) {
    return p.b_late && !p.b_ready && p.stage >= fwd_stages(p);
}

void add_stage(Iteration &it, Prod &p) {
    if (p.b_late) {
        const int nf = fwd_stages(p);
        if (p.b_ready && p.b_next < nf) add_fwd(it, p, p.b_next++, true);
        if (p.stage < nf) {
            add_fwd(it, p, p.stage++, false);
            return;
        }
        if (!p.b_ready || p.b_next < nf) return;  // B not yet caught up
        p.b_late = false;
    }
    for (int k = next_stage_span(p); k > 0; k--) add_one_stage(it, p);
}

void finalize_iteration(Iteration &it) {
    // transform items: group by (kind, radix)
    std::stable_sort(it.titems.begin(), it.titems.end(), [](const TItem &a, const TItem &b) {
        if (a.pad != b.pad) return a.pad < b.pad;
        return shape_key(a.d) < shape_key(b.d);
    });
    it.groups.clear();
    for (size_t i = 0; i < it.titems.size();) {
        size_t j = i;
        uint64_t tiles = 0;
        int lds_log = 0;
        while (j < it.titems.size() && it.titems[j].pad == it.titems[i].pad &&
               shape_key(it.titems[j].d) == shape_key(it.titems[i].d)) {
            const TItem &t = it.titems[j];
            it.titems[j].tile0 = tiles;
            // fused A+B items run on half-width tiles
            tiles += ((uint64_t)1 << t.log_n) >> (t.d.log_r + t.d.log_c - (t.pad == K_FUSED_AB ? 1 : 0));
            lds_log = std::max(lds_log, t.d.log_r + t.d.log_c);
            j++;
        }
        it.groups.push_back(Group{it.titems[i].pad, shape_key(it.titems[i].d), i, j - i, tiles, lds_log});
        i = j;
    }
    // carry items: group by step
    std::stable_sort(it.citems.begin(), it.citems.end(),
                     [](const CItem &a, const CItem &b) { return a.tile0 < b.tile0; });
    for (size_t i = 0; i < it.citems.size();) {
        size_t j = i;
        const uint64_t step = it.citems[i].tile0;
        uint64_t tiles = 0;
        while (j < it.citems.size() && it.citems[j].tile0 == step) j++;
        for (size_t q = i; q < j; q++) {
            it.citems[q].tile0 = tiles;
            tiles += step == 1 ? 1 : it.citems[q].ntiles;
        }
        it.groups.push_back(Group{K_CARRY + (int)step, 0, i, j - i, tiles, 0});
        i = j;
    }
}

hipError_t LaunchPacer::before() {
    while ((int)inflight.size() >= depth) {
        hipError_t e = hipEventSynchronize(inflight.front());
        if (e != hipSuccess) return e;
        spare.push_back(inflight.front());
        inflight.erase(inflight.begin());
    }
    return hipSuccess;
}

hipError_t LaunchPacer::after() {
    hipEvent_t e = nullptr;
    if (!spare.empty()) {
        e = spare.back();
        spare.pop_back();
    } else if (hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming)) {
        return r;
    }
    hipError_t r = hipEventRecord(e, st);
    if (r != hipSuccess) {
        spare.push_back(e);
        return r;
    }
    inflight.push_back(e);
    return hipSuccess;
}

void LaunchPacer::release() {
    for (hipEvent_t e : inflight) (void)hipEventDestroy(e);
    for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    inflight.clear();
    spare.clear();
}

template <int KIND>
static int launch_tgroup(const Group &g, const TItem *dev, const Tw &tw, hipStream_t st, LaunchPacer *pace) {
    const size_t lds = ((size_t)1 << g.lds_log) * 8;
    if (g.tiles >= ((uint64_t)1 << 31)) return -1;
    const TItem *items = dev + g.first;
    const int n = (int)g.count;
    const uint64_t piece = pace ? std::max<uint64_t>(pace->max_tiles, 1) : g.tiles;
    for (uint64_t base = 0; base < g.tiles; base += piece) {
        if (pace && pace->before() != hipSuccess) return -2;
        dim3 grid((unsigned)std::min<uint64_t>(piece, g.tiles - base));
        switch (g.shape) {
#define TP_CASE(LR, LC, S)                                                                                        \
    case shape_key(LR, LC, S):                                                                                    \
        hipLaunchKernelGGL((k_tpass<Geom<LR, LC, S>, KIND>), grid, dim3(TBLOCK),                                  \
                           (KIND == K_FUSED_AB ? fused_ab_lds_bytes<Geom<LR, LC, S>>() : lds), st, items, n, g.tiles,   \
                           base, tw);                                                                             \
        break;
            GNT_FOR_EACH_SHAPE(TP_CASE)
#undef TP_CASE
            default: return -1;
        }
        if (pace && pace->after() != hipSuccess) return -2;
    }
    return 0;
}

int launch_iteration(const Iteration &it, const TItem *dev_t, const CItem *dev_c, const Tw &tw, hipStream_t st,
                     LaunchPacer *pace) {
    for (const Group &g : it.groups) {
        int rc = 0;
        if (g.kind == K_FWD) rc = launch_tgroup<K_FWD>(g, dev_t, tw, st, pace);
        else if (g.kind == K_INV) rc = launch_tgroup<K_INV>(g, dev_t, tw, st, pace);
        else if (g.kind == K_INV_SUMS) rc = launch_tgroup<K_INV_SUMS>(g, dev_t, tw, st, pace);
        else if (g.kind == K_INV_PLAIN) rc = launch_tgroup<K_INV_PLAIN>(g, dev_t, tw, st, pace);
        else if (g.kind == K_INV_TSTORE) rc = launch_tgroup<K_INV_TSTORE>(g, dev_t, tw, st, pace);
        else if (g.kind == K_INV_SUMS_PLAIN) rc = launch_tgroup<K_INV_SUMS_PLAIN>(g, dev_t, tw, st, pace);
        else if (g.kind == K_FUSED) rc = launch_tgroup<K_FUSED>(g, dev_t, tw, st, pace);
        else if (g.kind == K_FUSED_AB) rc = launch_tgroup<K_FUSED_AB>(g, dev_t, tw, st, pace);
        else {
            // carry steps: short grid-stride launches, paced whole
            if (pace && pace->before() != hipSuccess) return -2;
            const CItem *items = dev_c + g.first;
            const int n = (int)g.count;
            dim3 grid((unsigned)std::min<uint64_t>(g.tiles, 65536));
            if (g.kind == K_CARRY) hipLaunchKernelGGL(k_ctiles, grid, dim3(CARRY_BLOCK), 0, st, items, n, g.tiles);
            else if (g.kind == K_CARRY + 1) hipLaunchKernelGGL(k_cscan, dim3((unsigned)n), dim3(SCAN_BLOCK), 0, st, items);
            else {
                hipLaunchKernelGGL(k_cripple, grid, dim3(CARRY_RIPPLE_BLOCK), 0, st, items, n, g.tiles);
                if (it.want_lens)
                    hipLaunchKernelGGL(k_lens, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, items, n);
            }
            if (pace && pace->after() != hipSuccess) return -2;
        }
        if (rc) return rc;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace gnt
