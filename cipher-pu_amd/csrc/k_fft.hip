// Canonical-embedding encode/decode (the only floating-point code on the
// path).  Slot j sits at zeta^(5^j mod 2N), zeta = exp(i pi / N)
// (scripts/verify_multi_slot_simd.py:32-91 evaluates at the same odd powers
// of zeta, in natural order; the 5^j order makes automorphisms act as slot
// rotations).
//
// With n = N/2 and u_i = m_i + i m_(i+n), m(zeta^(4k+1)) = sum_i u_i zeta^i
// w^(ik), w = exp(2 pi i / n), so decode is a pre-twist, one length-n complex
// FFT and a gather k_j = (5^j mod 2N - 1) / 4; encode runs it backwards.
//
// Integer sides: encode rounds to int64 and reduces into every limb;
// decode composes the limbs exactly (multi-word CRT), centers the value and
// converts to double.
#include "launch.h"

namespace {

// One length-n complex FFT in LDS, radix-2 DIT on bit-reversed input.
// sign > 0 uses w^(+ik) (table tw), sign < 0 its conjugate.
__device__ void fft_lds(double2* s, const double2* __restrict__ tw, int logn, int sign) {
    const int n = 1 << logn;
    for (int h = 1; h < n; h <<= 1) {
        const int stride = n / (2 * h);
        for (int t = threadIdx.x; t < n / 2; t += blockDim.x) {
            const int k = t & (h - 1);
            const int i0 = ((t / h) * 2 * h) + k;
            double2 w = tw[k * stride];
            if (sign < 0) w.y = -w.y;
            const double2 u = s[i0], v0 = s[i0 + h];
            const double2 v = make_double2(v0.x * w.x - v0.y * w.y, v0.x * w.y + v0.y * w.x);
            s[i0] = make_double2(u.x + v.x, u.y + v.y);
            s[i0 + h] = make_double2(u.x - v.x, u.y - v.y);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int brev(int i, int bits) { return (int)(__brev((unsigned)i) >> (32 - bits)); }

// coeffs [vec][N] (real, already divided by the scale) -> slots [vec][n]
__global__ void k_decode_fft(const double2* __restrict__ tw, const double2* __restrict__ zeta,
                             const int* __restrict__ slot_k, const double* __restrict__ coeffs,
                             double2* __restrict__ slots, int logn) {
    extern __shared__ double2 s[];
    const int n = 1 << logn;
    const double* m = coeffs + (size_t)blockIdx.x * 2 * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double re = m[i], im = m[i + n];
        const double2 z = zeta[i];
        s[brev(i, logn)] = make_double2(re * z.x - im * z.y, re * z.y + im * z.x);
    }
    __syncthreads();
    fft_lds(s, tw, logn, +1);
    double2* o = slots + (size_t)blockIdx.x * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) o[j] = s[slot_k ? slot_k[j] : j];
}

// slots [vec][n] (already multiplied by the scale) -> coeffs [vec][N]
__global__ void k_encode_fft(const double2* __restrict__ tw, const double2* __restrict__ zeta,
                             const int* __restrict__ slot_k, const double2* __restrict__ slots,
                             double* __restrict__ coeffs, int logn) {
    extern __shared__ double2 s[];
    const int n = 1 << logn;
    const double2* z = slots + (size_t)blockIdx.x * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) s[brev(slot_k ? slot_k[j] : j, logn)] = z[j];
    __syncthreads();
    fft_lds(s, tw, logn, -1);
    const double inv_n = 1.0 / n;
    double* m = coeffs + (size_t)blockIdx.x * 2 * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double2 y = s[i], c = zeta[i];   // times conj(zeta^i)
        m[i] = (y.x * c.x + y.y * c.y) * inv_n;
        m[i + n] = (y.y * c.x - y.x * c.y) * inv_n;
    }
}

// round to nearest (ties to even) and reduce into every limb
__global__ void k_encode_round(const double* __restrict__ coeffs, uint64_t* __restrict__ pt,
                               const PrimeConst* __restrict__ pcs, LimbMap map, int l, int logN, size_t total,
                               double scale) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const size_t N = (size_t)1 << logN;
    const size_t n = g & (N - 1), vec = g >> logN;
    const double c = rint(coeffs[g] * scale);
    const bool neg = c < 0;
    const uint64_t a = (uint64_t)(neg ? -c : c);
    uint64_t* o = pt + vec * l * N + n;
    for (int i = 0; i < l; ++i) {
        const PrimeConst& pc = pcs[map.idx[i]];
        const uint64_t r = reduce64(a, pc);
        o[(size_t)i * N] = (neg && r) ? pc.q - r : r;
    }
}

__device__ __forceinline__ bool mw_geq(const uint64_t* a, const uint64_t* b, int w) {
    for (int k = w - 1; k >= 0; --k)
        if (a[k] != b[k]) return a[k] > b[k];
    return true;
}

__device__ __forceinline__ void mw_sub(uint64_t* a, const uint64_t* b, int w) {
    uint64_t borrow = 0;
    for (int k = 0; k < w; ++k) {
        const uint64_t bk = b[k] + borrow;
        const uint64_t nb = (bk < borrow) || (a[k] < bk);
        a[k] -= bk;
        borrow = nb;
    }
}

// Exact CRT: x = sum_i [c_i Qhat_i^-1]_(q_i) Qhat_i mod Q, centered, as
// double, times inv_scale.
__global__ void k_decode_crt(const uint64_t* __restrict__ pt, double* __restrict__ coeffs,
                             const PrimeConst* __restrict__ pcs, LimbMap map, CrtConst cc, int logN, size_t total,
                             double inv_scale) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int l = cc.l, W = cc.words;
    const size_t N = (size_t)1 << logN;
    const size_t n = g & (N - 1), vec = g >> logN;
    const uint64_t* src = pt + vec * l * N + n;
    uint64_t acc[CRT_MAX_WORDS];
    for (int k = 0; k < W; ++k) acc[k] = 0;
    for (int i = 0; i < l; ++i) {
        const PrimeConst& pc = pcs[map.idx[i]];
        const uint64_t y = csub(mul_shoup(src[(size_t)i * N], cc.qhat_inv[i], cc.qhat_inv_s[i], pc.q), pc.q);
        // acc += y * Qhat_i
        uint64_t carry = 0;
        for (int k = 0; k < W; ++k) {
            const uint64_t w = cc.qhat[i][k];
            const uint64_t lo = y * w, hi = __umul64hi(y, w);
            uint64_t s = acc[k] + lo;
            uint64_t c = s < lo;
            s += carry;
            c += s < carry;
            acc[k] = s;
            carry = hi + c;
        }
    }
    for (int r = 0; r < l && mw_geq(acc, cc.Q, W); ++r) mw_sub(acc, cc.Q, W);
    bool neg = false;
    if (!mw_geq(cc.half_Q, acc, W)) {   // acc > Q/2: acc - Q < 0, magnitude Q - acc
        uint64_t t[CRT_MAX_WORDS];
        for (int k = 0; k < W; ++k) t[k] = cc.Q[k];
        mw_sub(t, acc, W);
        for (int k = 0; k < W; ++k) acc[k] = t[k];
        neg = true;
    }
    double d = 0.0;
    for (int k = W - 1; k >= 0; --k) d = d * 18446744073709551616.0 + (double)acc[k];
    coeffs[g] = (neg ? -d : d) * inv_scale;
}

}  // namespace

hipError_t launch_decode_fft(const double2* tw, const double2* zeta, const int* slot_k, const double* coeffs,
                             double2* slots, int n_vec, int logN, hipStream_t s) {
    if (!n_vec) return hipSuccess;
    const int logn = logN - 1, n = 1 << logn;
    const int thr = n / 2 < 64 ? 64 : (n / 2 > 512 ? 512 : n / 2);
    hipError_t e = hipFuncSetAttribute((const void*)k_decode_fft, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       n * 16);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_decode_fft, dim3(n_vec), dim3(thr), n * 16, s, tw, zeta, slot_k, coeffs, slots, logn);
    return hipGetLastError();
}

hipError_t launch_encode_fft(const double2* tw, const double2* zeta, const int* slot_k, const double2* slots,
                             double* coeffs, int n_vec, int logN, hipStream_t s) {
    if (!n_vec) return hipSuccess;
    const int logn = logN - 1, n = 1 << logn;
    const int thr = n / 2 < 64 ? 64 : (n / 2 > 512 ? 512 : n / 2);
    hipError_t e = hipFuncSetAttribute((const void*)k_encode_fft, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       n * 16);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_encode_fft, dim3(n_vec), dim3(thr), n * 16, s, tw, zeta, slot_k, slots, coeffs, logn);
    return hipGetLastError();
}

hipError_t launch_encode_round(const DevTables& t, const double* coeffs, uint64_t* pt, int n_vec, int l,
                               const LimbMap& map, double scale, hipStream_t s) {
    const size_t total = (size_t)n_vec << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_encode_round, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, coeffs, pt, t.pc,
                       map, l, t.logN, total, scale);
    return hipGetLastError();
}

hipError_t launch_decode_crt(const DevTables& t, const uint64_t* pt, double* coeffs, int n_vec, const LimbMap& map,
                             const CrtConst& cc, double inv_scale, hipStream_t s) {
    const size_t total = (size_t)n_vec << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_decode_crt, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, pt, coeffs, t.pc, map,
                       cc, t.logN, total, inv_scale);
    return hipGetLastError();
}
