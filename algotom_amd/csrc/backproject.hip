// Linear-interpolation back-projection on gfx950.
//
// Reference arithmetic (algotom/rec/reconstruction.py:156-190, 229-264,
// 307-359), per pixel (x, y) and angle i, with numba's typing:
//   c, s  = float32 cos/sin of float32 -angle[i]
//   f     = ((x - ic) * c + (y - ic) * s) + center        (float64; ic =
//           0.5 (ncol - 1), both products exact in float64)
//   if 0 <= f <= ncol-1:  d = floor f,  val = s[d] + (s[d+1] - s[d]) * (f - d)
//                         (float32 difference, float64 product and sum)
//   accumulate val: float64 register (2D kernel) or float32 after every
//   angle (chunk and CPU kernels).
//
// Kernel design (bp_tiled_kernel): a 256-thread workgroup owns a 64 x (4*PY)
// pixel tile of SB slices.  Lane l of wave w owns pixel column x0+l and the
// PY rows y0 + w*PY + k, so one wave instruction touches 64 consecutive x
// and its LDS gathers hit neighbouring or shared detector columns.  Angles
// are processed in blocks of A: for each angle the workgroup computes the
// tile's detector footprint [dlo, dlo+SPAN) exactly from its four corners
// and stages one record per (column, slice) into LDS -- the sample and its
// forward difference, interleaved over the SB slices so one pixel reads
// all of them with one or two ds_read_b128 (RecPlanes below).  f, d and
// the weight are computed once per pixel and angle and reused by the SB
// slices.  Angles
// whose footprint lies inside [0, ncol-1] for the whole tile (every interior
// tile) skip the range test; the rest take the exact per-pixel test.  Each
// pixel sums its angles in ascending order, as the reference does, so
// ATH_ACC_F64 and ATH_ACC_F32SEQ reproduce the reference kernels bit for bit
// given the same filtered sinogram.
//
// Per-mode arithmetic (Mode<> below):
//   ATH_ACC_F64     record float32 (s, ds); val and sum in float64
//   ATH_ACC_F32SEQ  record float32 (s, ds); val in float64, sum stored float32
//   ATH_ACC_F32     record float32 (s, ds); one FMA and one add in float32
//   kAccH16/kAccB16 float16/bfloat16 storage: record (s, ds) as a 16-bit pair,
//                   weight pair (1, w) with w rounded to nearest; one
//                   v_dot2c_f32_{f16,bf16} per update accumulating in float32.
#include "ath_internal.h"

namespace ath {
namespace {

constexpr int kThreads = 256;
constexpr int kTileX = 64;
constexpr int kAccH16 = 16;  // internal modes: 16-bit storage + dot2
constexpr int kAccB16 = 17;
constexpr int kAccF64W = 18;  // ATH_ACC_F64 on single slices, float64 records
constexpr int kAccF32SeqW = 19;  // ATH_ACC_F32SEQ likewise
constexpr int kAccDerivW = 20;     // ATH_ACC_DERIV likewise
// The single-slice float64-record form of a mode, or the mode itself.
// Measured on single 1801-angle slices: f64 2560 columns 244.7 -> 237.7 ms
// for 64 slices, f32seq 282.0 -> 239.6 ms, the centre derivative 3.23 ->
// 2.86 ms (22.4 -> 19.9 ms at 7192 columns).  Image + derivative in one
// pass (ATH_ACC_F64_DERIV) keeps float32 records: 4.91 against 5.25 ms.
constexpr int wide1_of(int mode) {
    return mode == ATH_ACC_F64 ? kAccF64W
         : mode == ATH_ACC_F32SEQ ? kAccF32SeqW
         : mode == ATH_ACC_DERIV ? kAccDerivW
         : mode;
}
constexpr bool is_wide1(int mode) {
    return mode == kAccF64W || mode == kAccF32SeqW || mode == kAccDerivW;
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf162_t __attribute__((ext_vector_type(2)));
typedef float float2_t __attribute__((ext_vector_type(2)));

constexpr int isqrt_ceil(int v) {
    int r = 0;
    while (r * r < v) ++r;
    return r;
}
// Columns staged per angle: the footprint of a TX x TY tile spans at most
// ceil(sqrt((TX-1)^2 + (TY-1)^2)) + 1 columns (|cos|, |sin| float32 with
// cos^2 + sin^2 <= 1 + 2^-23); two more for rounding slack.
constexpr int bp_span(int tx, int ty) {
    return isqrt_ceil((tx - 1) * (tx - 1) + (ty - 1) * (ty - 1)) + 3;
}
// Modes that evaluate f in float32 stage from floor(fmin) - 1 and one more
// column, so float32 rounding just outside the exact range still lands on a
// staged (zero-padded) record and interpolates continuously.
template <class M>
constexpr int mode_span(int tx, int ty) {
    return bp_span(tx, ty) + (M::kF32Pos ? 2 : 0);
}

// floor(f) as an int in one instruction (f within int range); with
// __builtin_amdgcn_fractf(f) == f - floorf(f) (exact for |f| < 2^23) it
// replaces floor, convert and subtract
__device__ __forceinline__ int floor_to_int(float f) {
    int d;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(d) : "v"(f));
    return d;
}

// detector position exactly as the reference evaluates it
__device__ __forceinline__ double det_pos(double xc_c, double yc, double s,
                                          double cen) {
    return fma(yc, s, xc_c) + cen;
}

__device__ __forceinline__ float load_f32(const float* p) { return *p; }
__device__ __forceinline__ float load_f32(const __half* p) {
    return __half2float(*p);
}
__device__ __forceinline__ float load_f32(const __hip_bfloat16* p) {
    return __bfloat162float(*p);
}

template <int MODE>
struct Mode;

template <>
struct Mode<ATH_ACC_F64> {
    static constexpr bool kF32Pos = false;
    static constexpr bool kTwoOut = false;
    static constexpr bool kPlanes = false;
    using Rec = float2;
    using W = double;
    using Acc = double;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        return make_float2(yd, yu - yd);
    }
    static __device__ __forceinline__ W weight(double fr) { return fr; }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W w) {
        a += (double)q.x + (double)q.y * w;
    }
    static __device__ __forceinline__ void add_value(Acc& a, float e) {
        a += (double)e;
    }
    static __device__ __forceinline__ float result(Acc a) { return (float)a; }
};

// ATH_ACC_F64 for one slice per workgroup with the record staged as the
// exactly converted float64 pair: the two converts per update move to the
// staging loop and the sums are bit-identical.  A 16-byte record is one
// ds_read_b128 per pixel (the float32 pair of 4 slices takes two).
template <>
struct Mode<kAccF64W> {
    static constexpr bool kF32Pos = false;
    static constexpr bool kTwoOut = false;
    static constexpr bool kPlanes = false;
    using Rec = double2;
    using W = double;
    using Acc = double;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        return make_double2((double)yd, (double)(yu - yd));
    }
    static __device__ __forceinline__ W weight(double fr) { return fr; }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W w) {
        a += q.x + q.y * w;
    }
    static __device__ __forceinline__ void add_value(Acc& a, float e) {
        a += (double)e;
    }
    static __device__ __forceinline__ float result(Acc a) { return (float)a; }
};

// ATH_ACC_F32SEQ for one slice per workgroup with float64 records, as
// kAccF64W: the same bits, the converts moved into staging.
template <>
struct Mode<kAccF32SeqW> {
    static constexpr bool kF32Pos = false;
    static constexpr bool kTwoOut = false;
    static constexpr bool kPlanes = false;
    using Rec = double2;
    using W = double;
    using Acc = float;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        return make_double2((double)yd, (double)(yu - yd));
    }
    static __device__ __forceinline__ W weight(double fr) { return fr; }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W w) {
        a = (float)((double)a + (q.x + q.y * w));
    }
    static __device__ __forceinline__ void add_value(Acc& a, float e) {
        a = (float)((double)a + (double)e);
    }
    static __device__ __forceinline__ float result(Acc a) { return a; }
};

template <>
struct Mode<ATH_ACC_F32SEQ> {
    static constexpr bool kF32Pos = false;
    static constexpr bool kTwoOut = false;
    static constexpr bool kPlanes = false;
    using Rec = float2;
    using W = double;
    using Acc = float;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        return make_float2(yd, yu - yd);
    }
    static __device__ __forceinline__ W weight(double fr) { return fr; }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W w) {
        a = (float)((double)a + ((double)q.x + (double)q.y * w));
    }
    static __device__ __forceinline__ void add_value(Acc& a, float e) {
        a = (float)((double)a + (double)e);
    }
    static __device__ __forceinline__ float result(Acc a) { return a; }
};

template <>
struct Mode<ATH_ACC_F32> {
    static constexpr bool kF32Pos = false;
    static constexpr bool kTwoOut = false;
    static constexpr bool kPlanes = true;
    using Rec = float2;
    using W = float;
    using Acc = float;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        return make_float2(yd, yu - yd);
    }
    static __device__ __forceinline__ W weight(double fr) { return (float)fr; }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W w) {
        a += fmaf(q.y, w, q.x);
    }
    static __device__ __forceinline__ void add_value(Acc& a, float e) {
        a += e;
    }
    static __device__ __forceinline__ float result(Acc a) { return a; }
};

// Derivative of the reference interpolation with respect to the centre:
// d/df [s[d] + (s[d+1] - s[d]) (f - d)] = s[d+1] - s[d] at d = floor f, and
// df/dcenter = 1.  Summed in float64; edge-padded samples do not move.
template <>
struct Mode<ATH_ACC_DERIV> {
    static constexpr bool kF32Pos = false;
    static constexpr bool kTwoOut = false;
    static constexpr bool kPlanes = false;
    using Rec = float2;
    using W = double;
    using Acc = double;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        return make_float2(yd, yu - yd);
    }
    static __device__ __forceinline__ W weight(double fr) { return fr; }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W) {
        a += (double)q.y;
    }
    static __device__ __forceinline__ void add_value(Acc&, float) {}
    static __device__ __forceinline__ float result(Acc a) { return (float)a; }
};

// ATH_ACC_F64 into recon and, from the same positions and records,
// ATH_ACC_DERIV into recon2: two float64 sums per pixel, bit-identical to
// the two modes run separately.
struct AccPair {
    double v, dv;
};

template <>
struct Mode<ATH_ACC_F64_DERIV> {
    static constexpr bool kF32Pos = false;
    static constexpr bool kTwoOut = true;
    static constexpr bool kPlanes = false;
    using Rec = float2;
    using W = double;
    using Acc = AccPair;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        return make_float2(yd, yu - yd);
    }
    static __device__ __forceinline__ W weight(double fr) { return fr; }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W w) {
        a.v += (double)q.x + (double)q.y * w;
        a.dv += (double)q.y;
    }
    static __device__ __forceinline__ void add_value(Acc& a, float e) {
        a.v += (double)e;
    }
    static __device__ __forceinline__ float result(Acc a) {
        return (float)a.v;
    }
    static __device__ __forceinline__ float result2(Acc a) {
        return (float)a.dv;
    }
};

// ATH_ACC_DERIV for one slice per workgroup with float64 records (as
// kAccF64W; the same bits).
template <>
struct Mode<kAccDerivW> {
    static constexpr bool kF32Pos = false;
    static constexpr bool kTwoOut = false;
    static constexpr bool kPlanes = false;
    using Rec = double2;
    using W = double;
    using Acc = double;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        return make_double2((double)yd, (double)(yu - yd));
    }
    static __device__ __forceinline__ W weight(double fr) { return fr; }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W) {
        a += q.y;
    }
    static __device__ __forceinline__ void add_value(Acc&, float) {}
    static __device__ __forceinline__ float result(Acc a) { return (float)a; }
};

template <>
struct Mode<kAccH16> {
    // detector position in float32 from a float64 per-angle offset
    static constexpr bool kF32Pos = true;
    static constexpr bool kTwoOut = false;
    static constexpr bool kPlanes = true;
    using Rec = half2_t;
    using W = half2_t;
    using Acc = float;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        // both samples are float16 values: the difference is exact in
        // float32 and rounds once to float16
        return Rec{(_Float16)yd, (_Float16)(yu - yd)};
    }
    static __device__ __forceinline__ W weight(double fr) {
        return weight32((float)fr);
    }
    static __device__ __forceinline__ W weight32(float fr) {
        // (1, w) in one v_cvt_pk_f16_f32: 1 is exact; w rounds to
        // nearest (rounding toward zero would pull every tap toward s[d]:
        // C1 error vs float64 2.9e-5 with nearest, 5.3e-5 toward zero)
        return __builtin_convertvector(float2_t{1.0f, fr}, W);
    }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W w) {
        a = __builtin_amdgcn_fdot2(q, w, a, false);
    }
    static __device__ __forceinline__ void add_value(Acc& a, float e) {
        a += e;
    }
    static __device__ __forceinline__ float result(Acc a) { return a; }
};

template <>
struct Mode<kAccB16> {
    // detector position in float32 from a float64 per-angle offset
    static constexpr bool kF32Pos = true;
    static constexpr bool kTwoOut = false;
    static constexpr bool kPlanes = true;
    using Rec = bf162_t;
    using W = bf162_t;
    using Acc = float;
    static __device__ __forceinline__ Rec make(float yd, float yu) {
        return Rec{(__bf16)yd, (__bf16)(yu - yd)};
    }
    static __device__ __forceinline__ W weight(double fr) {
        return weight32((float)fr);
    }
    static __device__ __forceinline__ W weight32(float fr) {
        // (1, w) in one v_cvt_pk_bf16_f32, w rounded to nearest; the
        // bfloat16 samples dominate the error (C1 vs float64: 2.71e-4 with
        // an exact weight, 2.82e-4 with this one)
        return __builtin_convertvector(float2_t{1.0f, fr}, W);
    }
    static __device__ __forceinline__ void add(Acc& a, Rec q, W w) {
        a = __builtin_amdgcn_fdot2_f32_bf16(q, w, a, false);
    }
    static __device__ __forceinline__ void add_value(Acc& a, float e) {
        a += e;
    }
    static __device__ __forceinline__ float result(Acc a) { return a; }
};

// mask (utility.py:160-182) and pi/(nrow-1) (reconstruction.py:493), in the
// float32 order numpy evaluates them
__device__ __forceinline__ float finalize(const BpArgs& a, float v, int x,
                                          int y) {
    if (a.use_mask) {
        const long long dx = x - a.mask_c, dy = y - a.mask_c;
        const float m = ((double)(dx * dx + dy * dy) <= a.mask_r2) ? 1.0f : 0.0f;
        v = v * m;
    }
    if (a.apply_scale) {
        v = v * a.pi32;
        v = v / a.denom;
    }
    return v;
}

// LDS record layout of one angle: SB slices per column, SPAN columns.
// By default a column's SB records are contiguous (one or two ds_read_b128
// per pixel).  Modes with kPlanes split them into NP planes of P slices (16
// bytes of records per column and plane), each plane SPAN columns long, so
// that neighbouring lanes on neighbouring columns read neighbouring 16-byte
// words.  Measured on MI355X with the 64 x 1801 x 2560 stack: the float32
// mode (4 slices, two planes) takes 105.6 ms in planes with 16-row tiles
// against 116.1 ms (32 rows) and 123.8 ms (16 rows) contiguous; the 16-bit
// modes (8 slices, two planes) 82.7 ms with 16 rows against 90.2 ms
// contiguous (and 102.6 ms in planes with 32 rows; padding the plane
// stride to other residues modulo 16 columns changed nothing).  The
// float64-sum modes are VALU-bound and 2-4 % slower in planes, so they
// stay contiguous.
template <class Rec, int SB, bool PLANES>
struct RecPlanes {
    static constexpr int kPer16 = (int)(16 / sizeof(Rec));
    static constexpr int P =
        PLANES && kPer16 >= 1 && kPer16 < SB ? kPer16 : SB;
    static constexpr int NP = SB / P;
    static_assert(NP * P == SB, "slice batch must be a multiple of the plane");
    // records from the column's first record to slice js's
    template <int SPAN>
    static constexpr int slice(int js) {
        return (js / P) * SPAN * P + js % P;
    }
};

// One angle over the thread's PY x SB pixels when the whole tile is in
// range: f >= 0, so the truncation is the floor and fract(f) = f - floor(f).
// The record address is one shift-add on a workgroup-uniform offset.
template <class M, int SB, int PY, int SPAN>
__device__ __forceinline__ void inner_angle(
    const typename M::Rec* s_rec, int al, double c, double s, int dlo,
    double xc, const double (&yc)[PY], double cen,
    typename M::Acc (&acc)[PY][SB]) {
    using Rec = typename M::Rec;
    using L = RecPlanes<Rec, SB, M::kPlanes>;
    // (a multiple of P records, so every plane read is one aligned
    // ds_read_b128)
    const int off =
        __builtin_amdgcn_readfirstlane(al * SPAN * L::NP - dlo) * L::P;
    const Rec* qbase = s_rec + off;
    if constexpr (M::kF32Pos) {
        // f = (y - ic) s + [(x - ic) c + center]; the bracket in float64,
        // rounded once; |error of f| < 3e-4 px at 8192 columns
        const float xcc = (float)(xc * c + cen);
        const float s32 = (float)s;
#pragma unroll
        for (int k = 0; k < PY; ++k) {
            const float f = fmaf((float)yc[k], s32, xcc);
            const int d = floor_to_int(f);
            const auto w = M::weight32(__builtin_amdgcn_fractf(f));
            const Rec* q = qbase + d * L::P;
#pragma unroll
            for (int js = 0; js < SB; ++js)
                M::add(acc[k][js], q[L::template slice<SPAN>(js)], w);
        }
        return;
    }
    const double xcc = xc * c;
#pragma unroll
    for (int k = 0; k < PY; ++k) {
        const double f = det_pos(xcc, yc[k], s, cen);
        const int d = (int)f;
        const auto w = M::weight(__builtin_amdgcn_fract(f));
        const Rec* q = qbase + d * L::P;
#pragma unroll
        for (int js = 0; js < SB; ++js)
            M::add(acc[k][js], q[L::template slice<SPAN>(js)], w);
    }
}

// Per-angle data of an angle block in two slots: the next block's is
// written while the current one's records are in use.
template <class Rec, int SB, int A, int SPAN>
struct alignas(16) BpShared {
    alignas(16) Rec rec[A * SPAN * SB];  // 16 B aligned: ds_read_b128 records
    double cos_[2][A], sin_[2][A];
    int dlo[2][A], inner[2][A];
    int all_inner[2];
    float edge[A * SB * 2];
};

// Angles b0 .. b0 + nb - 1 (nb <= A <= 64, so all in wave 0): float32
// cos / sin as the reference rounds them, the tile's detector footprint from
// its four corners, whether the footprint is inside [0, ncol-1], and
// whether that holds for every angle of the block.
template <class M, class Sh>
__device__ __forceinline__ void angle_setup(const BpArgs& a, Sh& sh, int sl,
                                            int b0, int nb, int tid,
                                            double cx0, double cx1,
                                            double cy0, double cy1,
                                            double cen, double w1) {
    if (tid >= 64) return;
    int inner = 1;
    if (tid < nb) {
        const float th = -a.angles[b0 + tid];
        const double c = (double)(float)cos((double)th);
        const double s = (double)(float)sin((double)th);
        const double f00 = det_pos(cx0 * c, cy0, s, cen);
        const double f01 = det_pos(cx1 * c, cy0, s, cen);
        const double f10 = det_pos(cx0 * c, cy1, s, cen);
        const double f11 = det_pos(cx1 * c, cy1, s, cen);
        const double fmn = fmin(fmin(f00, f01), fmin(f10, f11));
        const double fmx = fmax(fmax(f00, f01), fmax(f10, f11));
        sh.cos_[sl][tid] = c;
        sh.sin_[sl][tid] = s;
        sh.dlo[sl][tid] = (int)floor(fmn) - (M::kF32Pos ? 1 : 0);
        inner = (fmn >= 0.0) && (fmx <= w1);
        sh.inner[sl][tid] = inner;
    }
    const unsigned long long out = __ballot(!inner);
    if (tid == 0) sh.all_inner[sl] = out == 0ull ? 1 : 0;
}

// Back-projects tile (tx, ty) of slice group grp of one stack.  Called by
// every thread of the workgroup; starts and ends with all LDS consumed.
template <int MODE, int SB, int PY, int A, typename Tin>
__device__ __forceinline__ void bp_tile(
    const BpArgs& a, int tx, int ty, int grp,
    BpShared<typename Mode<MODE>::Rec, SB, A,
             mode_span<Mode<MODE>>(kTileX, 4 * PY)>& sh) {
    using M = Mode<MODE>;
    using Rec = typename M::Rec;
    using L = RecPlanes<Rec, SB, M::kPlanes>;
    static_assert(A <= 64, "angle_setup runs one block's angles in wave 0");
    constexpr int TY = 4 * PY;
    constexpr int SPAN = mode_span<M>(kTileX, TY);
    Rec* s_rec = sh.rec;
    float* s_edge = sh.edge;

    const Tin* __restrict__ sino = static_cast<const Tin*>(a.sino);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int x0 = tx * kTileX, y0 = ty * TY, j0 = grp * SB;
    const int C = a.ncol, S = a.nslice, N = a.nrow;
    const double ic = a.icenter, cen = a.center, w1 = (double)(C - 1);
    const int nsb = min(SB, S - j0);

    const int x = x0 + lane;
    const double xc = (double)x - ic;
    double yc[PY];
#pragma unroll
    for (int k = 0; k < PY; ++k) yc[k] = (double)(y0 + wv * PY + k) - ic;
    // corners of the whole tile (also covers lanes beyond the image edge)
    const double cx0 = (double)x0 - ic, cx1 = (double)(x0 + kTileX - 1) - ic;
    const double cy0 = (double)y0 - ic, cy1 = (double)(y0 + TY - 1) - ic;

    typename M::Acc acc[PY][SB];
#pragma unroll
    for (int k = 0; k < PY; ++k)
#pragma unroll
        for (int js = 0; js < SB; ++js) acc[k][js] = typename M::Acc{};

    // the first block's angles; each later block's are set up by wave 0
    // during the previous block's arithmetic, into the other slot, and
    // published by the barrier that ends that block
    angle_setup<M>(a, sh, 0, 0, min(A, N), tid, cx0, cx1, cy0, cy1, cen, w1);
    __syncthreads();
    for (int a0 = 0, sl = 0; a0 < N; a0 += A, sl ^= 1) {
        const int na = min(A, N - a0);
        const double* s_cos = sh.cos_[sl];
        const double* s_sin = sh.sin_[sl];
        const int* s_dlo = sh.dlo[sl];
        const int* s_inner = sh.inner[sl];
        const int all_inner = sh.all_inner[sl];
        // one record per (angle, column): the SB slices' (s, ds)
        const int nrec = na * SPAN;
        // two records per thread and step, both loaded before either is
        // stored; a record whose two columns are inside the row and whose
        // slice group is full loads without per-sample tests.  Against one
        // record per step with a test per sample: fp16 stack 117.4 ->
        // 90.7 ms, fp32 106.0 -> 103.6 ms, f32seq 185.9 -> 185.1 ms
        // (64 x 1801 x 2560)
        for (int e0 = tid; e0 < nrec; e0 += 2 * kThreads) {
            float yd[2][SB], yu[2][SB];
            int at[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = e0 + u * kThreads;
                const int al = e / SPAN;
                const int k = e - al * SPAN;
                at[u] = e < nrec ? al * SPAN * SB + k * L::P : -1;
#pragma unroll
                for (int js = 0; js < SB; ++js) {
                    yd[u][js] = 0.0f;
                    yu[u][js] = 0.0f;
                }
                if (e >= nrec) continue;
                const int col = s_dlo[al] + k;
                // column offsets are added only where they are in range
                const Tin* rows = sino + ((size_t)(a0 + al) * S + j0) * C;
                if (col >= 0 && col + 1 < C && nsb == SB) {
#pragma unroll
                    for (int js = 0; js < SB; ++js) {
                        const Tin* p = rows + (size_t)js * C + col;
                        yd[u][js] = load_f32(p);
                        yu[u][js] = load_f32(p + 1);
                    }
                } else {
                    const bool in0 = col >= 0 && col < C;
                    const bool in1 = col + 1 >= 0 && col + 1 < C;
#pragma unroll
                    for (int js = 0; js < SB; ++js) {
                        if (js < nsb) {
                            const Tin* row = rows + (size_t)js * C;
                            if (in0) yd[u][js] = load_f32(row + col);
                            if (in1) yu[u][js] = load_f32(row + (col + 1));
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (at[u] < 0) continue;
                Rec* dst = s_rec + at[u];
#pragma unroll
                for (int js = 0; js < SB; ++js)
                    dst[L::template slice<SPAN>(js)] =
                        M::make(yd[u][js], yu[u][js]);
            }
        }
        if (a.edge_pad) {
            for (int e = tid; e < na * SB; e += kThreads) {
                const int js = e % SB, al = e / SB;
                float e0 = 0.0f, e1 = 0.0f;
                if (js < nsb) {
                    const Tin* row =
                        sino + ((size_t)(a0 + al) * S + j0 + js) * C;
                    e0 = load_f32(row);
                    e1 = load_f32(row + C - 1);
                }
                s_edge[2 * e] = e0;
                s_edge[2 * e + 1] = e1;
            }
        }
        __syncthreads();
        if (a0 + A < N)
            angle_setup<M>(a, sh, sl ^ 1, a0 + A, min(A, N - a0 - A), tid,
                           cx0, cx1, cy0, cy1, cen, w1);

        if (all_inner) {
            // common case: the whole block is in range for the whole tile
            for (int al = 0; al < na; ++al)
                inner_angle<M, SB, PY, SPAN>(s_rec, al, s_cos[al], s_sin[al],
                                             s_dlo[al], xc, yc, cen, acc);
        } else {
            for (int al = 0; al < na; ++al) {
                if (s_inner[al]) {
                    inner_angle<M, SB, PY, SPAN>(s_rec, al, s_cos[al],
                                                 s_sin[al], s_dlo[al], xc, yc,
                                                 cen, acc);
                    continue;
                }
                const double c = s_cos[al], s = s_sin[al];
                const int dlo = s_dlo[al];
                const double xcc = xc * c;
                // float32-position modes interpolate exactly as inner_angle
                // does (so results do not depend on the tiling); the range
                // test is always the exact float64 one
                const float xcc32 = (float)(xc * c + cen);
                const float s32 = (float)s;
                const Rec* base = s_rec + al * SPAN * SB;
#pragma unroll
                for (int k = 0; k < PY; ++k) {
                    const double f = det_pos(xcc, yc[k], s, cen);
                    const bool in = (f >= 0.0) && (f <= w1);
                    int d;
                    typename M::W w;
                    if constexpr (M::kF32Pos) {
                        const float f32 = fmaf((float)yc[k], s32, xcc32);
                        const float fl = floorf(f32);
                        d = (int)fl;
                        w = M::weight32(f32 - fl);
                    } else {
                        d = (int)f;
                        w = M::weight(f - (double)d);
                    }
                    const int idx = min(max(d - dlo, 0), SPAN - 1);
                    const Rec* q = base + idx * L::P;
                    if (in) {
#pragma unroll
                        for (int js = 0; js < SB; ++js)
                            M::add(acc[k][js],
                                   q[L::template slice<SPAN>(js)], w);
                    } else if (a.edge_pad) {
                        const int side = f < 0.0 ? 0 : 1;
#pragma unroll
                        for (int js = 0; js < SB; ++js)
                            M::add_value(acc[k][js],
                                         s_edge[2 * (al * SB + js) + side]);
                    }
                }
            }
        }
        // the block's records and angles are consumed; the next block's
        // angles are published
        __syncthreads();
    }

    if (x < C) {
#pragma unroll
        for (int k = 0; k < PY; ++k) {
            const int y = y0 + wv * PY + k;
            if (y >= C) continue;
#pragma unroll
            for (int js = 0; js < SB; ++js) {
                if (js < nsb) {
                    const size_t o = ((size_t)y * S + j0 + js) * C + x;
                    a.recon[o] = finalize(a, M::result(acc[k][js]), x, y);
                    if constexpr (M::kTwoOut)
                        a.recon2[o] = finalize(a, M::result2(acc[k][js]), x,
                                               y);
                }
            }
        }
    }
}


template <int MODE, int SB, int PY, int A, typename Tin>
__global__ __launch_bounds__(kThreads) void bp_tiled_kernel(
    BpArgs a, int tiles_x, int tiles_y) {
    __shared__ BpShared<typename Mode<MODE>::Rec, SB, A,
                        mode_span<Mode<MODE>>(kTileX, 4 * PY)> sh;
    unsigned bid = blockIdx.x;
    const int tx = bid % tiles_x;
    bid /= tiles_x;
    const int ty = bid % tiles_y;
    const int grp = bid / tiles_y;
    bp_tile<MODE, SB, PY, A, Tin>(a, tx, ty, grp, sh);
}

// Persistent kernel over a work list of (stack, tile) items from many stacks
// of different widths, angle counts and centres.  Workgroups take the next
// item from a device counter (reset by the preceding reset_counter launch)
// until the list is exhausted; every workgroup reaches that exit.
template <int MODE, int SB, int PY, int A, typename Tin>
__global__ __launch_bounds__(kThreads) void bp_batched_kernel(
    const ath_bp_job* __restrict__ jobs, int njobs, long long total,
    int* __restrict__ counter) {
    __shared__ BpShared<typename Mode<MODE>::Rec, SB, A,
                        mode_span<Mode<MODE>>(kTileX, 4 * PY)> sh;
    __shared__ long long s_item;
    __shared__ int s_job;
    constexpr int TY = 4 * PY;
    for (;;) {
        if (threadIdx.x == 0) {
            const long long item = atomicAdd(counter, 1);
            int lo = 0, hi = njobs - 1;  // last job with tile_begin <= item
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (jobs[mid].tile_begin <= item) lo = mid;
                else hi = mid - 1;
            }
            s_item = item;
            s_job = lo;
        }
        __syncthreads();
        const long long item = s_item;
        const int j = s_job;
        __syncthreads();
        if (item >= total) break;
        const ath_bp_job jb = jobs[j];
        BpArgs a;
        a.sino = jb.sino;
        a.recon = jb.recon;
        a.recon2 = nullptr;
        a.angles = jb.angles;
        a.nrow = jb.nrow;
        a.nslice = jb.nslice;
        a.ncol = jb.ncol;
        a.center = (double)jb.center;
        a.icenter = 0.5 * (double)(jb.ncol - 1);
        a.edge_pad = jb.edge_pad;
        a.use_mask = jb.mask_radius >= 0.0;
        a.mask_r2 = jb.mask_radius * jb.mask_radius;
        a.mask_c = jb.ncol / 2;
        a.apply_scale = jb.apply_scale;
        a.pi32 = (float)M_PI;
        a.denom = (float)(jb.nrow - 1);
        const int tiles_x = (jb.ncol + kTileX - 1) / kTileX;
        const int tiles_y = (jb.ncol + TY - 1) / TY;
        long long local = item - jb.tile_begin;
        const int tx = (int)(local % tiles_x);
        local /= tiles_x;
        const int ty = (int)(local % tiles_y);
        const int grp = (int)(local / tiles_y);
        bp_tile<MODE, SB, PY, A, Tin>(a, tx, ty, grp, sh);
    }
}

__global__ void reset_counter_kernel(int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

// One thread per (x, y, slice), sinogram read straight from global memory.
template <int MODE>
__global__ void bp_naive_kernel(BpArgs a) {
    using M = Mode<MODE>;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int j = blockIdx.z;
    const int C = a.ncol, S = a.nslice;
    if (x >= C || y >= C) return;
    const float* sino = static_cast<const float*>(a.sino);
    const double ic = a.icenter, w1 = (double)(C - 1);
    const double xc = (double)x - ic, yc = (double)y - ic;
    typename M::Acc acc = 0;
    for (int i = 0; i < a.nrow; ++i) {
        const float th = -a.angles[i];
        const double c = (double)(float)cos((double)th);
        const double s = (double)(float)sin((double)th);
        const double f = det_pos(xc * c, yc, s, a.center);
        const float* row = sino + ((size_t)i * S + j) * C;
        if (f >= 0.0 && f <= w1) {
            const int d = (int)f;
            const float yd = row[d];
            const float yu = d + 1 < C ? row[d + 1] : 0.0f;
            M::add(acc, M::make(yd, yu), M::weight(f - (double)d));
        } else if (a.edge_pad) {
            M::add_value(acc, f < 0.0 ? row[0] : row[C - 1]);
        }
    }
    a.recon[((size_t)y * S + j) * C + x] = finalize(a, M::result(acc), x, y);
}

template <int MODE, int SB, int PY, int A, typename Tin>
int launch_tiled(const BpArgs& a, hipStream_t stream) {
    constexpr int TY = 4 * PY;
    const int tiles_x = (a.ncol + kTileX - 1) / kTileX;
    const int tiles_y = (a.ncol + TY - 1) / TY;
    const long long groups = (a.nslice + SB - 1) / SB;
    const long long blocks = (long long)tiles_x * tiles_y * groups;
    if (blocks > 0x7fffffffLL)
        return set_error(ATH_ERR_UNSUPPORTED, "back-projection grid too large");
    hipLaunchKernelGGL((bp_tiled_kernel<MODE, SB, PY, A, Tin>),
                       dim3((unsigned)blocks), dim3(kThreads), 0, stream, a,
                       tiles_x, tiles_y);
    return check_hip(hipGetLastError(), "bp_tiled_kernel launch");
}

template <int MODE, int SB, int PY, int A, typename Tin>
int launch_batched(const ath_bp_job* jobs, int njobs, long long total,
                   int nblocks, int* counter, hipStream_t stream) {
    auto kern = bp_batched_kernel<MODE, SB, PY, A, Tin>;
    if (nblocks <= 0) {
        // workgroups resident at once on the current device (cached per
        // kernel and device)
        const int st = resident_grid(reinterpret_cast<const void*>(kern),
                                     kThreads, 0, &nblocks);
        if (st != ATH_OK) return st;
    }
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream,
                       counter);
    int st = check_hip(hipGetLastError(), "reset_counter_kernel launch");
    if (st != ATH_OK) return st;
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(kThreads), 0, stream, jobs,
                       njobs, total, counter);
    return check_hip(hipGetLastError(), "bp_batched_kernel launch");
}

// Angles staged per LDS block: the largest of 64 / 32 / 16 whose shared
// arrays stay within 53 KB (three workgroups per CU).  Larger blocks mean
// fewer barriers and more staging loads in flight per phase.  The angle
// order, hence every result bit, does not depend on it.  Measured on MI355X:
//   one slice (SB = 1, 64 angles): the 256-sinogram uneven batch 299.3 /
//     222.9 / 211.9 ms (F64 / fp32 / fp16) against 323.7 / 246.6 / 234.9 ms
//     with 16 angles and 243.9 ms (fp16) with 128;
//   16-bit records, 4 slices x 32 rows (32 angles): the 64 x 1801 x 2560
//     fp16 stack 117.1 ms against 122.0 with 16 and 136.7 with 64;
//   image + centre derivative, 2 slices x 16 rows (32 angles): 4.97 against
//     5.22 ms at 16 x 901 x 1024, 447.6 against 478.4 ms at 16 x 1801 x 7192;
//   the centre derivative alone stays at 16 (3.35 against 3.46 ms with 32;
//     3.16 against 3.29 ms with the current staging, 308 against 316 ms at
//     16 x 1801 x 7192);
//   float32 and float64 records, 4 slices: 16 (32 would not fit).
template <int MODE, int SB, int PY>
constexpr int angle_block() {
    if (MODE == ATH_ACC_DERIV) return 16;
    using Rec = typename Mode<MODE>::Rec;
    constexpr int span = mode_span<Mode<MODE>>(kTileX, 4 * PY);
    // records, two slots of cos / sin / dlo / inner, edge values
    constexpr size_t per_angle =
        (size_t)span * SB * sizeof(Rec) + 48 + 8 * SB;
    return per_angle * 64 <= 53000 ? 64 : (per_angle * 32 <= 53000 ? 32 : 16);
}

template <int MODE, typename Tin>
int batched_config(const ath_bp_job* jobs, int njobs, long long total,
                   int sb, int py, int nblocks, int* counter, hipStream_t s) {
#define ATH_BPB_CASE(SB_, PY_)                                             \
    if (sb == SB_ && py == PY_)                                            \
        return launch_batched<MODE, SB_, PY_, angle_block<MODE, SB_, PY_>(), \
                              Tin>(jobs, njobs, total, nblocks, counter, s);
    // single slices of the float64-sum modes: float64 records
    if constexpr (wide1_of(MODE) != MODE) {
        if (sb == 1)
            return batched_config<wide1_of(MODE), Tin>(jobs, njobs, total, sb,
                                                       py, nblocks, counter,
                                                       s);
    }
    ATH_BPB_CASE(1, 4) ATH_BPB_CASE(1, 8) ATH_BPB_CASE(1, 16)
    if constexpr (!is_wide1(MODE)) {
        ATH_BPB_CASE(4, 4) ATH_BPB_CASE(4, 8)
    }
#undef ATH_BPB_CASE
    return set_error(ATH_ERR_UNSUPPORTED,
                     "batched back-projection: no kernel for slice_batch " +
                         std::to_string(sb) + ", tile_rows " +
                         std::to_string(4 * py));
}

// Tile configurations compiled per mode: slice batch SB and rows per lane
// PY (tile 64 x 4*PY).  16-bit records also allow SB = 8 (32 bytes a pixel,
// the same as SB = 4 float32 records).
template <int MODE, typename Tin>
int launch_config(const BpArgs& a, int sb, int py, hipStream_t s) {
    constexpr bool wide = (MODE == kAccH16 || MODE == kAccB16);
#define ATH_BP_CASE(SB_, PY_)                                              \
    if (sb == SB_ && py == PY_)                                            \
        return launch_tiled<MODE, SB_, PY_, angle_block<MODE, SB_, PY_>(),  \
                            Tin>(                                          \
            a, s);
    // single slices of the float64-sum modes: float64 records
    if constexpr (wide1_of(MODE) != MODE) {
        if (sb == 1) return launch_config<wide1_of(MODE), Tin>(a, sb, py, s);
    }
    ATH_BP_CASE(1, 4) ATH_BP_CASE(1, 8) ATH_BP_CASE(1, 16)
    if constexpr (!is_wide1(MODE)) {
        ATH_BP_CASE(2, 4) ATH_BP_CASE(2, 8) ATH_BP_CASE(2, 16)
        ATH_BP_CASE(4, 4) ATH_BP_CASE(4, 8) ATH_BP_CASE(4, 16)
    }
    if constexpr (wide) {
        ATH_BP_CASE(8, 4) ATH_BP_CASE(8, 8)
    }
#undef ATH_BP_CASE
    return set_error(ATH_ERR_UNSUPPORTED,
                     "back-projection: no kernel for slice_batch " +
                         std::to_string(sb) + ", tile_rows " +
                         std::to_string(4 * py));
}

// Default configuration per mode and stack depth (measured on MI355X with
// a 64 x 1801 x 2560 stack; the autotuner can override it per shape).
template <int MODE>
void default_config(int nslice, int& sb, int& py) {
    sb = nslice <= 2 ? nslice : 4;
    py = 8;
    // float64 sums: 4 x 16 (164.9 against 193.9 ms for 4 x 32); float32
    // records in planes: 4 x 16 (105.6 against 121.3 ms)
    if (nslice > 2 && (MODE == ATH_ACC_F64 || MODE == ATH_ACC_F64_DERIV ||
                       MODE == ATH_ACC_F32))
        py = 4;
    // 16-bit records, 8 slices in two planes x 16 rows: 82.7 against
    // 90.2 ms for 4 x 32
    if (nslice >= 8 && (MODE == kAccH16 || MODE == kAccB16)) {
        sb = 8;
        py = 4;
    }
    // the centre derivative reads one value per record: 2 slices x 16 rows
    // (16 x 901 x 1024: 3.47 ms against 4.84 ms for 4 x 32; 16 x 1801 x
    // 7192: 308 against 429 ms)
    if (MODE == ATH_ACC_DERIV && nslice >= 2) {
        sb = 2;
        py = 4;
    }
    // two float64 sums per pixel and slice: 2 x 16 keeps 120 VGPRs (4 x 16
    // takes 244)
    if (MODE == ATH_ACC_F64_DERIV && nslice >= 2) {
        sb = 2;
        py = 4;
    }
}

template <int MODE, typename Tin>
int dispatch(const BpArgs& a, int sb, int py, hipStream_t s) {
    if (sb <= 0 || py <= 0) default_config<MODE>(a.nslice, sb, py);
    return launch_config<MODE, Tin>(a, sb, py, s);
}

}  // namespace

int back_projection(const BpArgs& a, int sino_dtype, int acc, int sb,
                    int py, hipStream_t s) {
    if (acc == ATH_ACC_DERIV || acc == ATH_ACC_F64_DERIV) {
        if (sino_dtype != ATH_F32)
            return set_error(ATH_ERR_UNSUPPORTED,
                             "centre derivative needs a float32 sinogram");
        if (acc == ATH_ACC_DERIV)
            return dispatch<ATH_ACC_DERIV, float>(a, sb, py, s);
        if (!a.recon2)
            return set_error(ATH_ERR_INVALID_ARG,
                             "ATH_ACC_F64_DERIV needs the derivative output");
        return dispatch<ATH_ACC_F64_DERIV, float>(a, sb, py, s);
    }
    if (sino_dtype == ATH_F32) {
        switch (acc) {
        case ATH_ACC_F64: return dispatch<ATH_ACC_F64, float>(a, sb, py, s);
        case ATH_ACC_F32SEQ:
            return dispatch<ATH_ACC_F32SEQ, float>(a, sb, py, s);
        case ATH_ACC_F32: return dispatch<ATH_ACC_F32, float>(a, sb, py, s);
        }
        return set_error(ATH_ERR_INVALID_ARG, "back-projection: bad acc mode");
    }
    if (acc != ATH_ACC_F32)
        return set_error(ATH_ERR_UNSUPPORTED,
                         "back-projection: float16/bfloat16 sinograms need "
                         "acc = ATH_ACC_F32");
    if (sino_dtype == ATH_F16) return dispatch<kAccH16, __half>(a, sb, py, s);
    if (sino_dtype == ATH_BF16)
        return dispatch<kAccB16, __hip_bfloat16>(a, sb, py, s);
    return set_error(ATH_ERR_INVALID_ARG, "back-projection: bad sinogram dtype");
}

int back_projection_batched(const ath_bp_job* jobs, int njobs,
                            long long total, int sino_dtype, int acc, int sb,
                            int py, int nblocks, int* counter, hipStream_t s) {
    if (sino_dtype == ATH_F32) {
        switch (acc) {
        case ATH_ACC_F64:
            return batched_config<ATH_ACC_F64, float>(jobs, njobs, total, sb,
                                                      py, nblocks, counter, s);
        case ATH_ACC_F32SEQ:
            return batched_config<ATH_ACC_F32SEQ, float>(
                jobs, njobs, total, sb, py, nblocks, counter, s);
        case ATH_ACC_F32:
            return batched_config<ATH_ACC_F32, float>(jobs, njobs, total, sb,
                                                      py, nblocks, counter, s);
        }
        return set_error(ATH_ERR_INVALID_ARG, "back-projection: bad acc mode");
    }
    if (acc != ATH_ACC_F32)
        return set_error(ATH_ERR_UNSUPPORTED,
                         "back-projection: float16/bfloat16 sinograms need "
                         "acc = ATH_ACC_F32");
    if (sino_dtype == ATH_F16)
        return batched_config<kAccH16, __half>(jobs, njobs, total, sb, py,
                                               nblocks, counter, s);
    if (sino_dtype == ATH_BF16)
        return batched_config<kAccB16, __hip_bfloat16>(
            jobs, njobs, total, sb, py, nblocks, counter, s);
    return set_error(ATH_ERR_INVALID_ARG, "back-projection: bad sinogram dtype");
}

int back_projection_naive(const BpArgs& a, int acc, hipStream_t s) {
    const dim3 block(16, 16);
    const dim3 grid((a.ncol + 15) / 16, (a.ncol + 15) / 16, a.nslice);
    switch (acc) {
    case ATH_ACC_F64:
        hipLaunchKernelGGL(bp_naive_kernel<ATH_ACC_F64>, grid, block, 0, s, a);
        break;
    case ATH_ACC_F32SEQ:
        hipLaunchKernelGGL(bp_naive_kernel<ATH_ACC_F32SEQ>, grid, block, 0, s,
                           a);
        break;
    case ATH_ACC_F32:
        hipLaunchKernelGGL(bp_naive_kernel<ATH_ACC_F32>, grid, block, 0, s, a);
        break;
    default:
        return set_error(ATH_ERR_INVALID_ARG, "back-projection: bad acc mode");
    }
    return check_hip(hipGetLastError(), "bp_naive_kernel launch");
}

}  // namespace ath
