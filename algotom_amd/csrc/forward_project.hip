// Adjoint of the back-projection: a forward projection (Radon transform with
// the back-projector's own linear-interpolation weights), for the backward
// pass of differentiable FBP.
//
// The back-projector adds, for pixel (x, y) and angle i with
// f = ((x - ic) c + (y - ic) s) + center in [0, ncol-1], d = floor f,
// t = f - d, the value (1 - t) s[d] + t s[d+1] (reconstruction.py:170-181).
// Its transpose sends pixel value g(x, y) to bin b with the hat weight
// w_b(f) = max(0, 1 - |f - b|) (that is 1 - t for b = d and t for b = d + 1),
// restricted to 0 <= f <= ncol - 1.  One thread owns one (angle, bin) pair
// and walks the image along the axis closer to the detector normal (rows
// when |c| >= |s|, so |slope| >= ~0.69): on each line the pixels with
// f in (b - 1, b + 1) are the K = ceil(2 / |slope|) (2 or 3) integers after
// the line's crossing of f = b - sign(slope), found with one multiply; each
// candidate costs its hat weight and one float32 FMA per slice, with no
// branch (a candidate just outside the interval gets weight 0).  Sums run in
// float32 over two chunks of kRows lines and in float64 across those; the
// order per bin is fixed (no atomics), so results are deterministic.
#include "ath_internal.h"

namespace ath {
namespace {

// A workgroup = 4 waves = 4 consecutive angles x the same 64 detector bins x
// FSB slices.  The image is walked in chunks of kRows lines along the major
// axis; for each chunk the union of the four angles' windows over bins
// [d0 - 2, d0 + 65] is staged into LDS once (line-major, the FSB slices of a
// pixel in one record of RS floats; columns outside the image or the mask
// staged as zeros, so candidates there add nothing) and every lane gathers
// its candidates from LDS.  Each staged pixel serves up to 4 angles x 2
// bins.  Chunks whose window is wider than kMaxMinor columns, and angles
// with |slope| < 2 / kMaxK, read global memory instead (exact windows, the
// same hat weights).  Records are padded to RS = FSB + 4 floats for FSB 8
// and 16 so that lanes on neighbouring pixels start on different 16-byte
// bank groups.
// Walking rows, a chunk's window is contiguous in memory.  Walking columns
// (x = line), it is four floats wide per image row; with a transposed copy
// of the image (image_t) those blocks stage contiguous runs as well.
// Without it, at 16 x 7192^2, their 16-byte pieces of 128-byte lines hit
// the L2 27 % of the time, 4.9 TB came from beyond it per launch and the
// kernel took 873 against 516 ms.
// staged columns per line: the window of 68 bins at |slope| >= 0.69 is
// about 100 columns, plus the spread of the four angles' windows, which
// grows with the distance from the centre (up to about 27 columns at 3600
// pixels and 0.1 degree steps: 3 % of the chunks of a 7192-wide image then
// read global memory).  160 columns stage those too but measured slower
// (548.6 against 516.1 ms at 16 x 7192^2, 5.59 against 5.41 ms at
// 16 x 1024^2): the staging loop takes three steps instead of two.  A
// 120-column tile (39 KB, four workgroups per CU, VGPRs capped at 128 with
// spills) measured slower too: 6.52 against 5.99 ms at 1024 columns.
constexpr int kMaxMinor = 128;
constexpr int kMaxK = 8;
// below this |slope| a line runs (almost) along the ray: take it whole
constexpr double kMinSlope = 1.0 / 1024.0;

template <int FSB> struct FpRec { static constexpr int rs = FSB; };
template <> struct FpRec<8> { static constexpr int rs = 12; };
template <> struct FpRec<16> { static constexpr int rs = 20; };

// acc[j] += w(t) * q[j] for the K candidates of one (bin, line) run:
// t = f - bin of the first candidate, consecutive candidates RS floats apart
// in LDS and `slope` apart in t.  EDGE: the block holds bin 0 or the last
// bin, whose weights also need t in [t_lo, t_hi] (0 <= f <= ncol - 1).
template <int FSB, int RS, int K, bool EDGE>
__device__ __forceinline__ void hat_run(const float* __restrict__ q, double t,
                                        double slope, double t_lo,
                                        double t_hi, float (&acc)[FSB]) {
#pragma unroll
    for (int n = 0; n < K; ++n) {
        double w = fmax(1.0 - fabs(t), 0.0);
        if constexpr (EDGE) w = (t >= t_lo && t <= t_hi) ? w : 0.0;
        const float w32 = (float)w;
        const float* r = q + n * RS;
        // the last of three or more candidates is outside (d-1, d+1) for
        // most lanes when |slope| is near 1: those lanes skip its record
        // (the same bits; less LDS traffic: 5.81 against 6.05 ms)
        const bool skip = n == K - 1 && K >= 3 && w32 == 0.0f;
        if (!skip) {
            if constexpr (FSB % 4 == 0) {
#pragma unroll
                for (int h = 0; h < FSB / 4; ++h) {
                    const float4 v = reinterpret_cast<const float4*>(r)[h];
                    acc[4 * h] = fmaf(w32, v.x, acc[4 * h]);
                    acc[4 * h + 1] = fmaf(w32, v.y, acc[4 * h + 1]);
                    acc[4 * h + 2] = fmaf(w32, v.z, acc[4 * h + 2]);
                    acc[4 * h + 3] = fmaf(w32, v.w, acc[4 * h + 3]);
                }
            } else {
#pragma unroll
                for (int j = 0; j < FSB; ++j)
                    acc[j] = fmaf(w32, r[j], acc[j]);
            }
        }
        t += slope;
    }
}

// The lines m0..m1 of a staged chunk for one (angle, bin) lane with K
// candidates per line (K = 0: the runtime length kdyn).  Along the walking
// axis everything is linear in m: the crossing u(m) = ic + (dedge -
// base(m)) / slope steps by du per line, the run starts at kf = floor(u) + 1
// and its first t = f(kf) - d = (kf - u) slope + (dedge - d).  Returns the
// lanes (ballot bits) whose run left the staged range on some line: those
// runs read the zero records, and the caller redoes the lanes' lines from
// global memory.
template <int FSB, int RS, int K, bool EDGE>
__device__ __forceinline__ unsigned long long run_lines(
    const float* tile, const float* zpad, int m0, int m1, int ka, int wid,
    int kdyn, double u, double du, double slope, double dedge_d,
    double t_lo, double t_hi, float (&acc)[FSB]) {
    const int kk = K > 0 ? K : kdyn;
    const unsigned span = (unsigned)(wid - kk);
    unsigned long long bad = 0;
    for (int m = m0; m <= m1; ++m) {
        const int kf = (int)floor(u) + 1;
        const int k0 = kf - ka;
        // the window's one-column margins keep the run inside the staged
        // range; the test guards the LDS reads
        const bool ok = (unsigned)k0 <= span;
        bad |= __ballot(!ok);
        const float* q =
            ok ? tile + __umul24((unsigned)((m - m0) * kMaxMinor + k0), RS)
               : zpad;
        const double t = fma((double)kf - u, slope, dedge_d);
        if constexpr (K > 0) {
            hat_run<FSB, RS, K, EDGE>(q, t, slope, t_lo, t_hi, acc);
        } else {
            double tn = t;
            for (int n = 0; n < kdyn; ++n) {
                hat_run<FSB, RS, 1, EDGE>(q + n * RS, tn, slope, t_lo, t_hi,
                                          acc);
                tn += slope;
            }
        }
        u += du;
    }
    return bad;
}

// run_lines with K fixed at compile time for the common K = 2, 3 (|slope|
// >= 2/3) and 4; K is wave-uniform, so the branch is too.
template <int FSB, int RS, bool EDGE>
__device__ __forceinline__ unsigned long long run_lines_k(
    int K, const float* tile, const float* zpad, int m0, int m1, int ka,
    int wid, double u, double du, double slope, double dedge_d, double t_lo,
    double t_hi, float (&acc)[FSB]) {
#define ATH_RUN(KK)                                                          \
    run_lines<FSB, RS, KK, EDGE>(tile, zpad, m0, m1, ka, wid, K, u, du,      \
                                 slope, dedge_d, t_lo, t_hi, acc)
    if (K == 3) return ATH_RUN(3);
    if (K == 2) return ATH_RUN(2);
    if (K == 4) return ATH_RUN(4);
    return ATH_RUN(0);
#undef ATH_RUN
}

// Every pixel of line m whose f lies in (d - 1, d + 1), read from global
// memory with the exact window, the mask and the 0 <= f <= ncol - 1 test.
template <int FSB>
__device__ void gather_line_global(const FpArgs& a, int m, bool by_rows,
                                   bool steep, double slope, double inv_slope,
                                   double base, double dd, double t_lo,
                                   double t_hi, int j0, int ns,
                                   float (&acc)[FSB]) {
    const int C = a.ncol, S = a.nslice;
    const double ic = a.icenter;
    int k0 = 0, k1 = C - 1;
    if (steep) {
        double u0 = (dd - 1.0 - base) * inv_slope;
        double u1 = (dd + 1.0 - base) * inv_slope;
        if (u0 > u1) {
            const double t = u0;
            u0 = u1;
            u1 = t;
        }
        k0 = max((int)ceil(u0 + ic - 1e-6), 0);
        k1 = min((int)floor(u1 + ic + 1e-6), C - 1);
    }
    for (int k = k0; k <= k1; ++k) {
        const double t = ((double)k - ic) * slope + base - dd;
        const double w = 1.0 - fabs(t);
        if (!(w > 0.0 && t >= t_lo && t <= t_hi)) continue;
        const int x = by_rows ? k : m;
        const int y = by_rows ? m : k;
        if (a.use_mask) {
            const long long dx = x - a.mask_c, dy = y - a.mask_c;
            if (!((double)(dx * dx + dy * dy) <= a.mask_r2)) continue;
        }
        const float w32 = (float)w;
        const float* px = a.image + ((size_t)y * S + j0) * C + x;
#pragma unroll
        for (int j = 0; j < FSB; ++j)
            if (j < ns) acc[j] = fmaf(w32, px[(size_t)j * C], acc[j]);
    }
}

// True when no pixel of lines m0..m1 and minor columns ka..kb is inside
// both the image and the mask: with lines at distance >= dm from the mask
// centre, a kept pixel has |k - mc| <= sqrt(r2 - dm^2).  Two columns of
// slack cover runs that leave the staged range by a rounding step (those
// are redone from global memory, where the mask decides again).
__device__ __forceinline__ bool chunk_is_empty(const FpArgs& a, int m0,
                                               int m1, int ka, int kb) {
    const int C = a.ncol;
    if (kb < -2 || ka > C + 1) return true;
    if (!a.use_mask) return false;
    const int mc = a.mask_c;
    const int dm = mc < m0 ? m0 - mc : (mc > m1 ? mc - m1 : 0);
    const double rest = a.mask_r2 - (double)dm * (double)dm;
    if (rest < 0.0) return true;
    const double h = sqrt(rest) + 2.0;
    return (double)kb < (double)mc - h || (double)ka > (double)mc + h;
}

// Workgroups are dealt to the 8 XCDs round-robin by linear id, and each
// XCD has its own L2.  The blocks that read the same image columns are the
// consecutive angle blocks of one bin block, so the launch is 1-D and the
// id is remapped: XCD x runs a contiguous range of the order (angle block
// fastest, then bin block, then slice block).  Without it, a width of
// 113 bin blocks (7192 columns) put neighbouring angle blocks on different
// XCDs and every chunk window came from HBM.
struct FpBlock {
    int bin, ang, slc;  // bin block, angle block, slice block
};

template <int FSB>
__device__ __forceinline__ FpBlock fp_block(const FpArgs& a) {
    const int nab = (a.nrow + 3) / 4, nbb = (a.ncol + 63) / 64;
    const int nsb = (a.nslice + FSB - 1) / FSB;
    const int total = nab * nbb * nsb;
    const int p = blockIdx.x, x = p & 7, slot = p >> 3;
    const int per = total >> 3, rem = total & 7;
    const int l = x * per + min(x, rem) + slot;
    FpBlock b;
    b.ang = l % nab;
    const int rest = l / nab;
    b.bin = rest % nbb;
    b.slc = rest / nbb;
    return b;
}

// The column window of chunk lines m0..m1: the union over the workgroup's 4
// angles of the minor coordinates whose f lies in bins [d0 - 2, d0 + 65] on
// lines m0 and m1 (windows are linear in m).  Lanes 0-15 of wave 0 each
// evaluate one end, reduce min / max across those 16 lanes, and lane 0
// stores the staged column range and the skip flag into ab.
__device__ __forceinline__ void chunk_window(const FpArgs& a, int tid, int m0,
                                             int m1, int d0, bool by_rows,
                                             const double (&cs)[4][2],
                                             const double (&inv)[4],
                                             int* ab) {
    if (tid >= 64) return;
    const double ic = a.icenter, cen = a.center;
    double lo = 1e300, hi = -1e300;
    if (tid < 16) {
        const int q = tid >> 2, e = (tid >> 1) & 1, b = tid & 1;
        const double mc = (double)(e ? m1 : m0) - ic;
        const double base = (by_rows ? mc * cs[q][1] : mc * cs[q][0]) + cen;
        const double bin = b ? (double)d0 + 65.0 : (double)d0 - 2.0;
        // a near-parallel line (inverse 0) takes the whole line
        const double u = inv[q] != 0.0 ? (bin - base) * inv[q]
                                       : (b ? 1e300 : -1e300);
        lo = u;
        hi = u;
    }
    // lanes 0-15 are closed under these xor offsets
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off));
        hi = fmax(hi, __shfl_xor(hi, off));
    }
    // a negative inverse swaps the two bin ends: min / max over all sixteen
    // covers both orders.  The range is not clipped to the image: columns
    // outside it are staged as zeros.
    if (tid == 0) {
        const int C = a.ncol;
        const bool whole =
            lo <= -1e299 || hi >= 1e299 || hi - lo > 2.0 * kMaxMinor;
        const int ka = whole ? 0 : (int)floor(lo + ic) - 1;
        const int kb = whole ? C - 1 + kMaxMinor : (int)ceil(hi + ic) + 1;
        ab[0] = ka;
        ab[1] = kb;
        ab[2] = chunk_is_empty(a, m0, m1, ka, kb) ? 1 : 0;
    }
}

// Stages lines m0 .. m0 + nrows - 1 and minor columns ka .. ka + wid - 1
// of the chunk into the tile: the thread's kSteps pixels (mask and bounds
// test once per pixel, pixels outside staged as zeros) are loaded before any
// is stored, so all their loads are in flight together.  FULL: the slice
// group holds FSB slices (no per-slice test).  The former loop, one pixel
// per step with a test per slice, waited for each step's loads before the
// next: 5.45 against 4.91 ms at 16 x 901 x 1024, 519 against 459 ms at
// 16 x 1801 x 7192.
template <int FSB, int kRows, int RS, bool FULL>
__device__ __forceinline__ void stage_chunk(const FpArgs& a, float* tile,
                                            int tid, bool rowwise,
                                            bool by_rows, int m0, int nrows,
                                            int ka, int wid, int j0,
                                            int ns) {
    constexpr int kSteps = kRows * kMaxMinor / 256;
    static_assert(kSteps * 256 == kRows * kMaxMinor, "whole steps");
    const int C = a.ncol, S = a.nslice;
    float v[kSteps][FSB];
    int dst[kSteps];
#pragma unroll
    for (int it = 0; it < kSteps; ++it) {
        const int e = tid + it * 256;
        int r, k;
        if (rowwise) {
            r = e / kMaxMinor;
            k = e % kMaxMinor;
        } else {  // lanes along the lines (x) first
            r = e % kRows;
            k = e / kRows;
        }
        const bool in = r < nrows && k < wid;
        dst[it] = in ? (r * kMaxMinor + k) * RS : -1;
        const int kk = ka + k;
        const int x = by_rows ? kk : m0 + r;
        const int y = by_rows ? m0 + r : kk;
        bool keep = in && kk >= 0 && kk < C;
        if (keep && a.use_mask) {
            // integer squares and their sum are exact in float64
            const double dx = x - a.mask_c, dy = y - a.mask_c;
            keep = fma(dx, dx, dy * dy) <= a.mask_r2;
        }
        // zeros first, then the loads under the branch: 5 % faster than
        // zeros written in a separate branch of the store
#pragma unroll
        for (int js = 0; js < FSB; ++js) v[it][js] = 0.0f;
        if (keep) {
            const float* src =
                by_rows || !a.image_t
                    ? a.image + ((size_t)y * S + j0) * C + x
                    : a.image_t + ((size_t)x * S + j0) * C + y;
#pragma unroll
            for (int js = 0; js < FSB; ++js)
                if (FULL || js < ns) v[it][js] = src[(size_t)js * C];
        }
    }
#pragma unroll
    for (int it = 0; it < kSteps; ++it) {
        if (dst[it] < 0) continue;
        float* d = tile + dst[it];
        if constexpr (FSB % 4 == 0) {
#pragma unroll
            for (int q = 0; q < FSB / 4; ++q)
                reinterpret_cast<float4*>(d)[q] =
                    make_float4(v[it][4 * q], v[it][4 * q + 1],
                                v[it][4 * q + 2], v[it][4 * q + 3]);
        } else {
#pragma unroll
            for (int js = 0; js < FSB; ++js) d[js] = v[it][js];
        }
    }
}

template <int FSB, int kRows>
__global__ __launch_bounds__(256) void forward_project_hat_kernel(
    FpArgs a, int by_rows_block) {
    constexpr int RS = FpRec<FSB>::rs;
    // the tile, then kMaxK zero records: a run that would leave the staged
    // range reads those instead and its line is redone from global memory
    __shared__ __align__(16) float tile[(kRows * kMaxMinor + kMaxK) * RS];
    __shared__ double s_cs[4][2];
    __shared__ double s_inv[4];   // 1 / slope, 0 for near-parallel lines
    __shared__ int s_ab[2][3];    // chunk column range and skip, 2 slots
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int C = a.ncol, S = a.nslice, N = a.nrow;
    const FpBlock blk = fp_block<FSB>(a);
    const int d0 = blk.bin * 64;
    const int i = blk.ang * 4 + wv;
    const int j0 = blk.slc * FSB;
    const int ns = min(FSB, S - j0);
    const int d = d0 + lane;
    const double ic = a.icenter, cen = a.center;
    float* const zpad = tile + kRows * kMaxMinor * RS;
    for (int e = tid; e < kMaxK * RS; e += 256) zpad[e] = 0.0f;
    if (tid < 4) {
        const int ia = min(blk.ang * 4 + tid, N - 1);
        const float th = -a.angles[ia];
        s_cs[tid][0] = (double)(float)cos((double)th);
        s_cs[tid][1] = (double)(float)sin((double)th);
    }
    __syncthreads();
    const bool valid = (i < N) && (d < C);
    const double c = s_cs[wv][0], s = s_cs[wv][1];
    // walking direction of the whole block: rows when the first angle has
    // |cos| >= |sin| (exact windows keep any direction correct)
    const bool by_rows = by_rows_block < 0
                             ? fabs(s_cs[0][0]) >= fabs(s_cs[0][1])
                             : by_rows_block != 0;
    if (tid < 4) {
        const double sl = by_rows ? s_cs[tid][0] : s_cs[tid][1];
        s_inv[tid] = fabs(sl) >= kMinSlope ? 1.0 / sl : 0.0;
    }
    const bool rowwise = by_rows || a.image_t != nullptr;
    // f = (k - ic) * slope + base(m) along a line m of the walking axis
    const double slope = by_rows ? c : s;
    const double other = by_rows ? s : c;
    const bool steep = fabs(slope) >= kMinSlope;
    const double inv_slope = steep ? 1.0 / slope : 0.0;
    // candidates per line (wave-uniform): the integers in an open interval
    // of length 2 / |slope|
    const int K = __builtin_amdgcn_readfirstlane(
        steep ? (int)ceil(2.0 * fabs(inv_slope)) : kMaxK + 1);
    const bool fast_wave = K <= kMaxK;
    // the crossing f = d - sign(slope) starts the candidate run
    const double dd = (double)d;
    const double dedge = slope > 0.0 ? dd - 1.0 : dd + 1.0;
    // 0 <= f <= ncol - 1 matters only for bins 0 and ncol - 1 (elsewhere a
    // positive weight already puts f inside): t = f - d must stay >= 0 for
    // bin 0 and <= 0 for the last bin
    const bool edge_block = __builtin_amdgcn_readfirstlane(
                                (d0 == 0 || d0 + 64 > C - 1) ? 1 : 0) != 0;
    const double t_lo = d == 0 ? 0.0 : -2.0;
    const double t_hi = d == C - 1 ? 0.0 : 2.0;
    float acc32[FSB];
    double acc[FSB];
#pragma unroll
    for (int j = 0; j < FSB; ++j) {
        acc32[j] = 0.0f;
        acc[j] = 0.0;
    }

    // the first chunk's window; each later one is computed by wave 0
    // during the previous chunk's arithmetic into the other slot, and
    // published by the barrier that ends that chunk
    chunk_window(a, tid, 0, min(kRows, C) - 1, d0, by_rows, s_cs, s_inv,
                 s_ab[0]);
    __syncthreads();
    for (int m0 = 0, slot = 0; m0 < C; m0 += kRows, slot ^= 1) {
        const int m1 = min(m0 + kRows, C) - 1;
        const int ka = s_ab[slot][0], kb = s_ab[slot][1];
        // every pixel the chunk's runs (and their margins) can reach is
        // outside the mask or the image: its terms are all w * 0, so the
        // chunk is skipped (the flush schedule below is unchanged)
        const bool empty = s_ab[slot][2] != 0;
        const int wid = kb - ka + 1;
        const bool staged = wid <= kMaxMinor;
        if (staged && !empty) {
            // lanes run along the staged columns when those are contiguous
            // in memory: rows of the image, or columns of its transposed
            // copy
            const int nrows = m1 - m0 + 1;
            if (ns == FSB)
                stage_chunk<FSB, kRows, RS, true>(a, tile, tid, rowwise,
                                                  by_rows, m0, nrows, ka, wid,
                                                  j0, ns);
            else
                stage_chunk<FSB, kRows, RS, false>(a, tile, tid, rowwise,
                                                   by_rows, m0, nrows, ka,
                                                   wid, j0, ns);
        }
        __syncthreads();
        if (m0 + kRows < C)
            chunk_window(a, tid, m0 + kRows, min(m0 + 2 * kRows, C) - 1, d0,
                         by_rows, s_cs, s_inv, s_ab[slot ^ 1]);
        if (valid && !empty) {
            if (staged && fast_wave && wid >= K) {
                const double u0 =
                    (dedge - (((double)m0 - ic) * other + cen)) * inv_slope +
                    ic;
                const double du = -other * inv_slope;
                const unsigned long long bad =
                    edge_block
                        ? run_lines_k<FSB, RS, true>(
                              K, tile, zpad, m0, m1, ka, wid, u0, du, slope,
                              dedge - dd, t_lo, t_hi, acc32)
                        : run_lines_k<FSB, RS, false>(
                              K, tile, zpad, m0, m1, ka, wid, u0, du, slope,
                              dedge - dd, t_lo, t_hi, acc32);
                if ((bad >> lane) & 1ull) {
                    // this lane's runs that left the staged range (the
                    // same test as in run_lines), from global memory
                    double u = u0;
                    for (int m = m0; m <= m1; ++m, u += du) {
                        const int k0 = (int)floor(u) + 1 - ka;
                        if ((unsigned)k0 <= (unsigned)(wid - K)) continue;
                        gather_line_global<FSB>(
                            a, m, by_rows, steep, slope, inv_slope,
                            ((double)m - ic) * other + cen, dd, t_lo, t_hi,
                            j0, ns, acc32);
                    }
                }
            } else {
                // global memory, one line at a time (a near-parallel line
                // holds many candidates: fold per line)
                for (int m = m0; m <= m1; ++m) {
                    gather_line_global<FSB>(a, m, by_rows, steep, slope,
                                            inv_slope,
                                            ((double)m - ic) * other + cen,
                                            dd, t_lo, t_hi, j0, ns, acc32);
#pragma unroll
                    for (int j = 0; j < FSB; ++j) {
                        acc[j] += (double)acc32[j];
                        acc32[j] = 0.0f;
                    }
                }
            }
        }
        if (valid) {
            // float32 sums over two chunks (at most 2 * kRows * K terms),
            // then into float64
            if (((m0 / kRows) & 1) || m1 == C - 1) {
#pragma unroll
                for (int j = 0; j < FSB; ++j) {
                    acc[j] += (double)acc32[j];
                    acc32[j] = 0.0f;
                }
            }
        }
        __syncthreads();  // tile free for the next chunk
    }
    if (valid) {
#pragma unroll
        for (int j = 0; j < FSB; ++j)
            if (j < ns)
                a.sino[((size_t)i * S + j0 + j) * C + d] =
                    (float)(acc[j] * a.scale);
    }
}

// out(x, j, y) = in(y, j, x): 64 x 64 tiles through LDS (a padded row
// keeps the transposed reads conflict-free), slices over grid.y.
__global__ __launch_bounds__(256) void transpose_slices_kernel(
    const float* __restrict__ in, float* __restrict__ out, int S, int C,
    int ntx) {
    __shared__ float t[64][65];
    const int x0 = (blockIdx.x % ntx) * 64, y0 = (blockIdx.x / ntx) * 64;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    for (int j = blockIdx.y; j < S; j += gridDim.y) {
        for (int r = ly; r < 64; r += 4) {
            const int y = y0 + r, x = x0 + lx;
            if (y < C && x < C) t[r][lx] = in[((size_t)y * S + j) * C + x];
        }
        __syncthreads();
        for (int r = ly; r < 64; r += 4) {
            const int x = x0 + r, y = y0 + lx;
            if (x < C && y < C) out[((size_t)x * S + j) * C + y] = t[lx][r];
        }
        __syncthreads();
    }
}

}  // namespace

int transpose_slices(const float* in, float* out, int nslice, int ncol,
                     hipStream_t stream) {
    const int ntx = (ncol + 63) / 64;
    const long long ntile = (long long)ntx * ntx;
    if (ntile > 0x7fffffffLL)
        return set_error(ATH_ERR_UNSUPPORTED, "transpose: image too large");
    const dim3 grid((unsigned)ntile, (unsigned)std::min(nslice, 65535));
    hipLaunchKernelGGL(transpose_slices_kernel, grid, dim3(256), 0, stream,
                       in, out, nslice, ncol, ntx);
    return check_hip(hipGetLastError(), "transpose_slices_kernel launch");
}

// Angle blocks of 4 share one walking direction: rows when the block's
// first angle has |cos| >= |sin|.  The candidate windows are exact for any
// direction, so a block straddling 45 degrees stays correct (its windows
// are at most 2/0.70 + 3 wide).  Launched per direction-homogeneous run of
// blocks: one launch for each maximal run.
int forward_projection(const FpArgs& a, hipStream_t stream) {
    // host copy of the walking direction per block needs the angles: they
    // live on the device, so the kernel decides per block instead
    // slices per workgroup and lines per chunk measured on MI355X
    // (16 x 901 x 1024: 16/4 9.9 ms, 8/4 11.2 ms, 16/8 12.7 ms)
    const int fsb = a.nslice >= 16 ? 16
                  : a.nslice >= 8 ? 8 : (a.nslice >= 4 ? 4 : 1);
    // one 1-D grid in the order fp_block decodes
    const long long nblk = (long long)((a.ncol + 63) / 64) *
                           ((a.nrow + 3) / 4) *
                           ((a.nslice + fsb - 1) / fsb);
    if (nblk > 0x7fffffffLL)
        return set_error(ATH_ERR_UNSUPPORTED,
                         "forward projection: too many workgroups");
    const dim3 grid((unsigned)nblk);
    if (fsb == 16)
        hipLaunchKernelGGL((forward_project_hat_kernel<16, 4>), grid,
                           dim3(256), 0, stream, a, -1);
    else if (fsb == 8)
        hipLaunchKernelGGL((forward_project_hat_kernel<8, 4>), grid,
                           dim3(256), 0, stream, a, -1);
    else if (fsb == 4)
        hipLaunchKernelGGL((forward_project_hat_kernel<4, 8>), grid,
                           dim3(256), 0, stream, a, -1);
    else
        hipLaunchKernelGGL((forward_project_hat_kernel<1, 16>), grid,
                           dim3(256), 0, stream, a, -1);
    return check_hip(hipGetLastError(), "forward_project_hat_kernel launch");
}

}  // namespace ath
