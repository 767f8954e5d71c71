// Trial sampling, syndromes and the stabilizer-residual test on the device.
//
// Sampling is bit-exact with the reference's host loop
// (jointbp_ets.cpp:5932-5946): per-trial seed splitmix64(base ^ idx*phi)
// (jointbp_ets.cpp:4401-4411), one std::mt19937_64 stream per trial, one
// uniform double per qubit in qubit order through libstdc++'s
// generate_canonical<double, 53> (one 64-bit draw d -> double(d) * 2^-64,
// values that round to 1.0 clamped to the largest double below 1), then the
// depolarizing thresholds in double precision.
//
// Seeding (312 dependent 64-bit steps per trial) runs first, one lane per
// trial, into a scratch array of initial states.  Then one wavefront per
// trial: the 312-word Mersenne-Twister state lives in LDS, the twist runs in
// three dependency phases across the 64 lanes, and each lane tempers and
// classifies five of every 312 outputs.  Error bits are set in LDS bit
// planes by atomic OR; syndrome bits are assembled with wave ballots.
//
// Classification works on the raw 64-bit draw d.  The map d -> u (round to
// double, scale by 2^-64, clamp below 1) is non-decreasing, and so are
// u -> (u - (1-p)) / p under IEEE subtraction and division by p > 0; each of
// the reference's three comparisons is therefore "d below a threshold".  The
// host finds the three thresholds by bisection over d with the same double
// arithmetic (first_draw_above below), so the device compares integers and the
// result is identical to evaluating the doubles draw by draw.
#include "jbp_common.h"

#include <algorithm>

namespace jbp {
namespace {

constexpr int kMtN = 312;
constexpr int kMtM = 156;
constexpr uint64_t kMtA = 0xb5026f5aa96619e9ull;
constexpr uint64_t kUpper = 0xffffffff80000000ull;  // upper 33 bits (r = 31)
constexpr uint64_t kLower = 0x000000007fffffffull;

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// The twist and the tempering on 32-bit halves: a ^ (b & K) is one
// three-input bit operation (table 0x78 over (a, b, K)), a 64-bit shift an
// align-bit of the halves.  Same bits as
//   y = (cur & kUpper) | (next & kLower),
//   far ^ (y >> 1) ^ ((y & 1) ? kMtA : 0)
// and the standard tempering (shifts 29, 17, 37, 43).
__device__ __forceinline__ uint32_t lo32(uint64_t x) { return static_cast<uint32_t>(x); }
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return static_cast<uint32_t>(x >> 32); }
__device__ __forceinline__ uint32_t xor_and(uint32_t a, uint32_t b, uint32_t k) {
    return __builtin_amdgcn_bitop3_b32(a, b, k, 0x78);  // a ^ (b & k)
}
__device__ __forceinline__ uint64_t twist_word(uint64_t cur, uint64_t next, uint64_t far) {
    static_assert(kUpper == 0xffffffff80000000ull && kLower == 0x7fffffffull, "MT19937-64 masks");
    const uint32_t ch = hi32(cur), yl = (lo32(cur) & 0x80000000u) | (lo32(next) & 0x7fffffffu);
    const uint32_t m = 0u - (lo32(next) & 1u);  // all ones when y is odd
    const uint32_t lo = xor_and(__builtin_amdgcn_alignbit(ch, yl, 1), m, lo32(kMtA)) ^ lo32(far);
    const uint32_t hi = xor_and(ch >> 1, m, hi32(kMtA)) ^ hi32(far);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t temper(uint64_t z) {
    const uint32_t hi = hi32(z), lo = lo32(z);
    const uint32_t lo1 = xor_and(lo, __builtin_amdgcn_alignbit(hi, lo, 29), 0x55555555u);  // >> 29
    const uint32_t hi1 = xor_and(hi, hi >> 29, 0x55555555u);
    const uint32_t lo2 = xor_and(lo1, lo1 << 17, 0xeda60000u);                            // << 17
    const uint32_t hi2 = xor_and(hi1, __builtin_amdgcn_alignbit(hi1, lo1, 15), 0x71d67fffu);
    const uint32_t hi3 = xor_and(hi2, lo2 << 5, 0xfff7eee0u);                             // << 37
    const uint32_t lo4 = lo2 ^ (hi3 >> 11);                                                // >> 43
    return (static_cast<uint64_t>(hi3) << 32) | lo4;
}

// Syndromes (jointbp_ets.cpp:1834-1849) of one trial from its error bit
// planes in LDS (wn + 1 words each, the last word 0: padded entries of the
// check table name variable n), written as bit planes; one wavefront, 64
// checks per ballot.  Each lane reads its check's variables as CW/4 aligned
// int4 loads, all in flight together.
template <int CW>
__device__ void write_syndromes(const DevGraph &g, TrialPlanes &tp, int trial, const uint32_t *ex,
                                const uint32_t *ez) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int m = g.side[s].m;
        const int4 *rows = reinterpret_cast<const int4 *>(g.cpad[s]);
        const uint32_t *bits = s == 0 ? ex : ez;
        uint32_t *out = (s == 0 ? tp.syn_x : tp.syn_z) + static_cast<size_t>(trial) * g.wm[s];
        for (int c0 = 0; c0 < m; c0 += 64) {
            const int c = c0 + lane;
            uint32_t par = 0;
            if (c < m) {
                int4 v[CW / 4];
#pragma unroll
                for (int k = 0; k < CW / 4; ++k) v[k] = rows[static_cast<size_t>(c) * (CW / 4) + k];
#pragma unroll
                for (int k = 0; k < CW / 4; ++k)
                    par ^= (bits[v[k].x >> 5] >> (v[k].x & 31)) ^ (bits[v[k].y >> 5] >> (v[k].y & 31)) ^
                           (bits[v[k].z >> 5] >> (v[k].z & 31)) ^ (bits[v[k].w >> 5] >> (v[k].w & 31));
            }
            const uint64_t mk = __ballot(par & 1u);
            if (lane < 2 && (c0 >> 5) + lane < g.wm[s])
                out[(c0 >> 5) + lane] = static_cast<uint32_t>(mk >> (32 * lane));
        }
    }
}

// Syndromes of caller-supplied error planes; grid = count, block = 64,
// dynamic LDS = 2*(wn+1) words.
template <int CW>
__global__ void __launch_bounds__(64) syndrome_kernel(DevGraph g, TrialPlanes tp) {
    extern __shared__ uint32_t planes[];
    const int trial = blockIdx.x;
    const int wn = g.wn;
    uint32_t *ex = planes;
    uint32_t *ez = planes + wn + 1;
    for (int w = threadIdx.x; w <= wn; w += 64) {
        ex[w] = w < wn ? tp.err_x[static_cast<size_t>(trial) * wn + w] : 0u;
        ez[w] = w < wn ? tp.err_z[static_cast<size_t>(trial) * wn + w] : 0u;
    }
    __syncthreads();
    write_syndromes<CW>(g, tp, trial, ex, ez);
}

// std::mt19937_64 seeding of trials first_trial + t, t < count: state[0] =
// splitmix64(base ^ idx*phi), state[i] = 6364136223846793005 *
// (state[i-1] ^ (state[i-1] >> 62)) + i, into rows of kMtN words.  The chain
// is serial within a trial, so each lane seeds its own trial.
__global__ void __launch_bounds__(64) seed_kernel(uint64_t base_seed, int64_t first_trial, int count,
                                                  uint64_t *state) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= count) return;
    const uint64_t idx = static_cast<uint64_t>(first_trial + t);
    uint64_t x = splitmix64(base_seed ^ (idx * 0x9e3779b97f4a7c15ull));
    uint64_t *row = state + static_cast<size_t>(t) * kMtN;
    row[0] = x;
    for (int i = 1; i < kMtN; ++i) {
        x = 6364136223846793005ull * (x ^ (x >> 62)) + static_cast<uint64_t>(i);
        row[i] = x;
    }
}

// Trials trial0 + blockIdx.x of the planes, from their seeded states
// (state row blockIdx.x); grid = trials, block = 64 (one wavefront),
// dynamic LDS = 2*(wn+1) words.
template <int CW>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_num_vgpr(16))) sample_kernel(DevGraph g, TrialPlanes tp, const uint64_t *state,
                                                    int trial0, NoiseTab nt) {
    __shared__ uint64_t mt[kMtN];
    extern __shared__ uint32_t planes[];  // err_x | err_z, wn + 1 words each
    const int lane = threadIdx.x;
    const int trial = trial0 + static_cast<int>(blockIdx.x);
    const int n = g.n, wn = g.wn;
    const DrawThresholds th = nt.th[nt.entry(trial)];
    uint32_t *ex = planes;
    uint32_t *ez = planes + wn + 1;
    for (int w = lane; w < 2 * (wn + 1); w += 64) planes[w] = 0u;
    const uint64_t *row = state + static_cast<size_t>(blockIdx.x) * kMtN;
    for (int k = lane; k < kMtN; k += 64) mt[k] = row[k];
    __syncthreads();

    for (int base = 0; base < n; base += kMtN) {
        // Twist phase A: k in [0, 156) reads old k, k+1, k+156.
        uint64_t nv[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            int k = lane + 64 * j;
            if (k < kMtM) nv[j] = twist_word(mt[k], mt[k + 1], mt[k + kMtM]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            int k = lane + 64 * j;
            if (k < kMtM) mt[k] = nv[j];
        }
        __syncthreads();
        // Phase B: k in [156, 311) reads old k, k+1 and new k-156.
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            int k = kMtM + lane + 64 * j;
            if (k < kMtN - 1) nv[j] = twist_word(mt[k], mt[k + 1], mt[k - kMtM]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            int k = kMtM + lane + 64 * j;
            if (k < kMtN - 1) mt[k] = nv[j];
        }
        __syncthreads();
        // Phase C: the last word reads new 0 and new 155.
        if (lane == 0) mt[kMtN - 1] = twist_word(mt[kMtN - 1], mt[0], mt[kMtM - 1]);
        __syncthreads();

        // each lane sets its own qubits' bits (LDS atomic OR, no return
        // value to wait for; most draws set nothing)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const int k = lane + 64 * j;
            const int q = base + k;
            if (k < kMtN && q < n) {
                const uint64_t z = temper(mt[k]);
                const bool b0 = (th.all_below & 1u) || z < th.t[0];
                const bool b1 = (th.all_below & 2u) || z < th.t[1];
                const bool b2 = (th.all_below & 4u) || z < th.t[2];
                const int s = b0 ? 0 : (b1 ? 1 : (b2 ? 2 : 3));
                const uint32_t bit = 1u << (q & 31);
                if (s & 1) atomicOr(&ex[q >> 5], bit);
                if (s >> 1) atomicOr(&ez[q >> 5], bit);
            }
        }
        __syncthreads();
    }

    write_syndromes<CW>(g, tp, trial, ex, ez);
    uint32_t *gx = tp.err_x + static_cast<size_t>(trial) * wn;
    uint32_t *gz = tp.err_z + static_cast<size_t>(trial) * wn;
    for (int w = lane; w < wn; w += 64) {
        gx[w] = ex[w];
        gz[w] = ez[w];
    }
}

// Stabilizer-residual test (jointbp_ets.cpp:887-920, 6199-6207): the X part
// of est xor err must lie in rowspace(Hz) and the Z part in rowspace(Hx).
// With a reduced row-echelon basis, v is in the span iff v equals the xor of
// the basis rows whose pivot bit is set in v.  One block per trial.
constexpr int kResBlock = 256;

__device__ bool in_span(const DevBasis &b, const uint32_t *d, uint32_t *acc, int wn) {
    const int tid = threadIdx.x;
    for (int w = tid; w < wn; w += kResBlock) acc[w] = 0u;
    __syncthreads();
    __shared__ int sel[kResBlock];
    for (int i0 = 0; i0 < b.rank; i0 += kResBlock) {
        int i = i0 + tid;
        int hit = 0;
        if (i < b.rank) {
            int col = b.pivot[i];
            hit = (d[col >> 5] >> (col & 31)) & 1;
        }
        sel[tid] = hit;
        __syncthreads();
        int lim = min(kResBlock, b.rank - i0);
        for (int k = 0; k < lim; ++k) {
            if (!sel[k]) continue;
            const uint32_t *row = b.rows + static_cast<size_t>(i0 + k) * b.words;
            for (int w = tid; w < wn; w += kResBlock) acc[w] ^= row[w];
        }
        __syncthreads();
    }
    int bad = 0;
    for (int w = tid; w < wn; w += kResBlock) bad |= (acc[w] != d[w]);
    return !__syncthreads_or(bad);
}

__global__ void __launch_bounds__(kResBlock) residual_kernel(DevGraph g, DevBasis bx, DevBasis bz,
                                                            TrialPlanes tp) {
    extern __shared__ uint32_t sm[];
    const int wn = g.wn;
    uint32_t *dx = sm;
    uint32_t *dz = sm + wn;
    uint32_t *acc = sm + 2 * wn;
    const int trial = blockIdx.x;
    const int tid = threadIdx.x;
    const size_t off = static_cast<size_t>(trial) * wn;
    int nz_x = 0, nz_z = 0;
    for (int w = tid; w < wn; w += kResBlock) {
        dx[w] = tp.err_x[off + w] ^ tp.est_x[off + w];
        dz[w] = tp.err_z[off + w] ^ tp.est_z[off + w];
        nz_x |= dx[w] != 0u;
        nz_z |= dz[w] != 0u;
    }
    int any_x = __syncthreads_or(nz_x);
    int any_z = __syncthreads_or(nz_z);
    bool ok = tp.syn_ok[trial] != 0;
    bool eq = true;
    if (ok && any_x) eq = in_span(bz, dx, acc, wn);
    if (ok && eq && any_z) eq = in_span(bx, dz, acc, wn);
    if (tid == 0) {
        tp.exact[trial] = (!any_x && !any_z) ? 1 : 0;
        tp.success[trial] = (ok && eq) ? 1 : 0;
    }
}

// One trial per block: the residual planes in LDS, then one row of logical
// operators per thread iteration.
__global__ void __launch_bounds__(kResBlock) logical_kernel(DevGraph g, DevLogicals lx,
                                                            DevLogicals lz, TrialPlanes tp) {
    extern __shared__ uint32_t sm[];
    const int wn = g.wn;
    uint32_t *dx = sm;
    uint32_t *dz = sm + wn;
    const int trial = blockIdx.x;
    const int tid = threadIdx.x;
    const size_t off = static_cast<size_t>(trial) * wn;
    int nz = 0;
    for (int w = tid; w < wn; w += kResBlock) {
        dx[w] = tp.err_x[off + w] ^ tp.est_x[off + w];
        dz[w] = tp.err_z[off + w] ^ tp.est_z[off + w];
        nz |= (dx[w] | dz[w]) != 0u;
    }
    const int any = __syncthreads_or(nz);
    const bool ok = tp.syn_ok[trial] != 0;
    int odd = 0;
    if (ok && any) {
        for (int64_t r = tid; r < lx.rows && !odd; r += kResBlock) {
            uint32_t par = 0;
            for (int64_t e = lx.off[r]; e < lx.off[r + 1]; ++e) {
                const int q = lx.qubit[e];
                par ^= dx[q >> 5] >> (q & 31);
            }
            odd |= par & 1u;
        }
        for (int64_t r = tid; r < lz.rows && !odd; r += kResBlock) {
            uint32_t par = 0;
            for (int64_t e = lz.off[r]; e < lz.off[r + 1]; ++e) {
                const int q = lz.qubit[e];
                par ^= dz[q >> 5] >> (q & 31);
            }
            odd |= par & 1u;
        }
    }
    const int bad = __syncthreads_or(odd);
    if (tid == 0) {
        tp.exact[trial] = any ? 0 : 1;
        tp.success[trial] = (ok && !bad) ? 1 : 0;
    }
}

// Sort key of each trial: its syndrome weight (both sides), scaled to 12 bits.
__global__ void __launch_bounds__(64) weight_kernel(DevGraph g, TrialPlanes tp, int shift, int *keys) {
    const int trial = blockIdx.x;
    const int lane = threadIdx.x;
    int w = 0;
    for (int i = lane; i < g.wm[0]; i += 64) w += __popc(tp.syn_x[static_cast<size_t>(trial) * g.wm[0] + i]);
    for (int i = lane; i < g.wm[1]; i += 64) w += __popc(tp.syn_z[static_cast<size_t>(trial) * g.wm[1] + i]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) w += __shfl_xor(w, m);
    if (lane == 0) keys[trial] = min(w >> shift, 4095);
}

// One block: counting sort of the keys, largest first.  A pipelined run
// orders chunk j+1 between the end of its sampling and the start of its
// decoder launch; four wavefronts (one per SIMD) with few registers, so the
// block fits beside a resident decoder workgroup.
constexpr int kOrderBlock = 256;
constexpr int kOrderKeys = 4096 / kOrderBlock;  // bins per thread in the scan
__global__ void __launch_bounds__(kOrderBlock) order_kernel(const int *keys, int count, int *order) {
    __shared__ int bins[4096];
    __shared__ int part[kOrderBlock];
    const int tid = threadIdx.x;
    for (int b = tid; b < 4096; b += kOrderBlock) bins[b] = 0;
    __syncthreads();
    for (int t = tid; t < count; t += kOrderBlock) atomicAdd(&bins[keys[t]], 1);
    __syncthreads();
    // start of key k in descending order = number of trials with key > k;
    // thread i owns keys 4095 - kOrderKeys*i down to 4096 - kOrderKeys*(i+1)
    int loc[kOrderKeys];
    int sum = 0;
#pragma unroll
    for (int j = 0; j < kOrderKeys; ++j) {
        loc[j] = sum;
        sum += bins[4095 - (kOrderKeys * tid + j)];
    }
    part[tid] = sum;
    __syncthreads();
    for (int d = 1; d < kOrderBlock; d <<= 1) {  // inclusive scan of the per-thread sums
        const int v = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    const int before = part[tid] - sum;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kOrderKeys; ++j) bins[4095 - (kOrderKeys * tid + j)] = before + loc[j];
    __syncthreads();
    for (int t = tid; t < count; t += kOrderBlock) order[atomicAdd(&bins[keys[t]], 1)] = t;
}

__global__ void __launch_bounds__(64) fill_i32_kernel(int *p, int v, int count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// out[t][q] = xbit | zbit << 1  (0 = I, 1 = X, 2 = Z, 3 = Y)
__global__ void unpack_states_kernel(const uint32_t *xb, const uint32_t *zb, int words, int n,
                                     int count, int *out) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<int64_t>(n) * count) return;
    int t = static_cast<int>(i / n);
    int q = static_cast<int>(i % n);
    size_t w = static_cast<size_t>(t) * words + (q >> 5);
    int x = (xb[w] >> (q & 31)) & 1;
    int z = (zb[w] >> (q & 31)) & 1;
    out[i] = x | (z << 1);
}

// Bit planes from int vectors (nonzero = 1), one word per thread.
__global__ void pack_bits_kernel(const int *vals, int len, int words, int count, uint32_t *out) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<int64_t>(words) * count) return;
    int t = static_cast<int>(i / words);
    int w = static_cast<int>(i % words);
    uint32_t acc = 0;
    for (int b = 0; b < 32; ++b) {
        int q = w * 32 + b;
        if (q < len && vals[static_cast<size_t>(t) * len + q] != 0) acc |= 1u << b;
    }
    out[i] = acc;
}

}  // namespace

namespace {

// libstdc++ generate_canonical<double, 53> of one mt19937_64 draw.
double canonical(uint64_t d) {
    double u = static_cast<double>(d) * 0x1p-64;
    return u >= 1.0 ? 0x1.fffffffffffffp-1 : u;
}

// The reference's classification of a uniform u (jointbp_ets.cpp:5936-5945),
// as the index of the first test that fails: 0 = I, 1 = X, 2 = Z, 3 = Y.
int classify(double u, double p) {
    volatile double omp = 1.0 - p;  // keep the host's double rounding explicit
    if (u < omp) return 0;
    volatile double v = (u - omp) / p;
    if (v < 1.0 / 3.0) return 1;
    if (v < 2.0 / 3.0) return 2;
    return 3;
}

// Smallest draw whose state is above k (i.e. >= k+1); classify(canonical(d))
// is non-decreasing in d.  Returns false when no draw gets there.
bool first_draw_above(int k, double p, uint64_t &out) {
    if (classify(canonical(~0ull), p) <= k) return false;
    uint64_t lo = 0, hi = ~0ull;  // answer in [lo, hi]
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (classify(canonical(mid), p) > k) hi = mid;
        else lo = mid + 1;
    }
    out = lo;
    return true;
}

}  // namespace

DrawThresholds draw_thresholds(double p) {
    DrawThresholds th;
    th.all_below = 0;
    for (int k = 0; k < 3; ++k) {
        if (!first_draw_above(k, p, th.t[k])) {
            th.t[k] = 0;
            th.all_below |= 1u << k;
        }
    }
    return th;
}

void launch_sample(const DevGraph &g, TrialPlanes tp, uint64_t base_seed, int64_t first_trial,
                   int count, const NoiseTab &nt, hipStream_t st, uint64_t *mt_state, int mt_cap) {
    if (count <= 0 || mt_cap <= 0) return;
    const size_t lds = static_cast<size_t>(2) * (g.wn + 1) * sizeof(uint32_t);
    // pieces of at most mt_cap trials: seed their states, then sample them
    for (int t0 = 0; t0 < count; t0 += mt_cap) {
        const int k = std::min(mt_cap, count - t0);
        hipLaunchKernelGGL(seed_kernel, dim3((k + 63) / 64), dim3(64), 0, st, base_seed, first_trial + t0,
                           k, mt_state);
        switch (g.cpad_w) {
            case 8: hipLaunchKernelGGL((sample_kernel<8>), dim3(k), dim3(64), lds, st, g, tp, mt_state, t0, nt); break;
            case 12: hipLaunchKernelGGL((sample_kernel<12>), dim3(k), dim3(64), lds, st, g, tp, mt_state, t0, nt); break;
            case 16: hipLaunchKernelGGL((sample_kernel<16>), dim3(k), dim3(64), lds, st, g, tp, mt_state, t0, nt); break;
            default: hipLaunchKernelGGL((sample_kernel<32>), dim3(k), dim3(64), lds, st, g, tp, mt_state, t0, nt); break;
        }
    }
}

void launch_syndromes(const DevGraph &g, TrialPlanes tp, int count, hipStream_t st) {
    if (count <= 0) return;
    const size_t lds = static_cast<size_t>(2) * (g.wn + 1) * sizeof(uint32_t);
    switch (g.cpad_w) {
        case 8: hipLaunchKernelGGL((syndrome_kernel<8>), dim3(count), dim3(64), lds, st, g, tp); break;
        case 12: hipLaunchKernelGGL((syndrome_kernel<12>), dim3(count), dim3(64), lds, st, g, tp); break;
        case 16: hipLaunchKernelGGL((syndrome_kernel<16>), dim3(count), dim3(64), lds, st, g, tp); break;
        default: hipLaunchKernelGGL((syndrome_kernel<32>), dim3(count), dim3(64), lds, st, g, tp); break;
    }
}

void launch_residual(const DevGraph &g, DevBasis bx, DevBasis bz, TrialPlanes tp, int count,
                     hipStream_t st) {
    if (count <= 0) return;
    size_t lds = static_cast<size_t>(3) * g.wn * sizeof(uint32_t);
    hipLaunchKernelGGL(residual_kernel, dim3(count), dim3(kResBlock), lds, st, g, bx, bz, tp);
}

void launch_logical(const DevGraph &g, DevLogicals lx, DevLogicals lz, TrialPlanes tp, int count,
                    hipStream_t st) {
    if (count <= 0) return;
    size_t lds = static_cast<size_t>(2) * g.wn * sizeof(uint32_t);
    hipLaunchKernelGGL(logical_kernel, dim3(count), dim3(kResBlock), lds, st, g, lx, lz, tp);
}

void launch_order_by_weight(const DevGraph &g, TrialPlanes tp, int count, int *keys, int *order,
                            hipStream_t st) {
    if (count <= 0) return;
    int shift = 0;
    while (((g.side[0].m + g.side[1].m) >> shift) > 4095) ++shift;
    hipLaunchKernelGGL(weight_kernel, dim3(count), dim3(64), 0, st, g, tp, shift, keys);
    hipLaunchKernelGGL(order_kernel, dim3(1), dim3(kOrderBlock), 0, st, keys, count, order);
}

void launch_fill_i32(int *ptr, int value, int count, hipStream_t st) {
    if (count <= 0) return;
    // one-wavefront blocks: a fill queued between decoder launches starts on
    // whatever compute unit has a wavefront slot free
    hipLaunchKernelGGL(fill_i32_kernel, dim3((count + 63) / 64), dim3(64), 0, st, ptr, value,
                       count);
}

void launch_unpack_states(const uint32_t *xbits, const uint32_t *zbits, int words, int n,
                          int count, int *out, hipStream_t st) {
    int64_t total = static_cast<int64_t>(n) * count;
    if (total <= 0) return;
    hipLaunchKernelGGL(unpack_states_kernel, dim3(static_cast<unsigned>((total + 255) / 256)),
                       dim3(256), 0, st, xbits, zbits, words, n, count, out);
}

void launch_pack_bits(const int *vals, int len, int words, int count, uint32_t *out,
                      hipStream_t st) {
    int64_t total = static_cast<int64_t>(words) * count;
    if (total <= 0) return;
    hipLaunchKernelGGL(pack_bits_kernel, dim3(static_cast<unsigned>((total + 255) / 256)),
                       dim3(256), 0, st, vals, len, words, count, out);
}

}  // namespace jbp
