// Direct bottom solve of the multigrid hierarchy by separation of variables
// (gfx950).
//
// On a level whose operator has a uniform A a, a uniform face coefficient b
// and side models without a second-cell term (ghost = dep0 x0: Neumann,
// order-2 Dirichlet, Robin; or periodic), the operator is a Kronecker sum
//     L = A a I + sum_d B b h_d That_d,   h_d = 1/dx_d^2,
// of one-dimensional matrices That_d = tridiag(-1, 2, -1) with the corner
// entries 2 - dep0 of the non-periodic sides (the ghost folded in) or the
// wrap-around entries of a periodic direction.  Each That_d is symmetric:
// That_d = Q_d diag(mu_d) Q_d^T with Q_d orthogonal, so
//     L^{-1} f = (Q_x (x) Q_y (x) Q_z) diag(1/(A a + sum_d B b h_d mu_d))
//                (Q_x^T (x) Q_y^T (x) Q_z^T) f,
// evaluated as one matrix product along each direction (forward with Q^T,
// the eigenvalue division fused into the last forward product, then back
// with Q).  A mode with a vanishing eigenvalue sum (a singular operator:
// no Dirichlet or Robin side and A a = 0) gets a zero coefficient.
//
// The eigenbases come from cyclic Jacobi rotations on the host, once per
// (extent, side models) and cached; the eigenvalues scale with B b h_d per
// component.  Cost per product: n_d multiply-adds per cell.
#ifndef VD_MG_DIRECT_H_
#define VD_MG_DIRECT_H_

#include <vector>
#include <map>
#include <tuple>
#include <cmath>
#include <cstring>
#include <algorithm>

namespace vd {

// Eigen-decomposition of a symmetric n x n matrix (row-major) by cyclic
// Jacobi rotations: A = Q diag(lam) Q^T, the columns of Q the eigenvectors.
inline void sym_eigen(int n, std::vector<double> A, std::vector<double>& Q, std::vector<double>& lam)
{
    Q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    double fro = 0.0;
    for (double v : A) fro += v * v;
    for (int sweep = 0; sweep < 100; ++sweep)
    {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
        if (!(off > 1e-32 * fro)) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q)
            {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                // rotation J (J_pp = J_qq = c, J_pq = s, J_qp = -s) with
                // (J^T A J)_pq = 0
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k)
                {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k)
                {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k)
                {
                    const double qkp = Q[(size_t)k * n + p], qkq = Q[(size_t)k * n + q];
                    Q[(size_t)k * n + p] = c * qkp - s * qkq;
                    Q[(size_t)k * n + q] = s * qkp + c * qkq;
                }
            }
    }
    lam.resize(n);
    for (int i = 0; i < n; ++i) lam[i] = A[(size_t)i * n + i];
}

// The one-dimensional matrix That of a direction of n cells: the operator
// row h b (2 x_i - x_{i+1} - x_{i-1}) divided by h b, with the ghost
// dep0 x0 of a non-periodic side folded into the end cell's diagonal.
inline std::vector<double> sep_matrix(int n, bool periodic, double dep0_lo, double dep0_hi)
{
    std::vector<double> T((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
    {
        T[(size_t)i * n + i] += 2.0;
        if (i > 0) T[(size_t)i * n + i - 1] -= 1.0;
        else if (periodic) T[(size_t)i * n + n - 1] -= 1.0;
        else T[(size_t)i * n + i] -= dep0_lo;
        if (i < n - 1) T[(size_t)i * n + i + 1] -= 1.0;
        else if (periodic) T[(size_t)i * n + 0] -= 1.0;
        else T[(size_t)i * n + i] -= dep0_hi;
    }
    return T;
}

struct SepBasis
{
    std::vector<double> Q, mu;
};

// Cached eigenbases keyed by (n, periodic, dep0 lo, dep0 hi).
inline const SepBasis& sep_basis(int n, bool periodic, double dep0_lo, double dep0_hi)
{
    static std::map<std::tuple<int, int, double, double>, SepBasis> cache;
    const auto key = std::make_tuple(n, periodic ? 1 : 0, periodic ? 0.0 : dep0_lo, periodic ? 0.0 : dep0_hi);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    SepBasis b;
    sym_eigen(n, sep_matrix(n, periodic, dep0_lo, dep0_hi), b.Q, b.mu);
    return cache.emplace(key, std::move(b)).first->second;
}

// One matrix product along a direction of a level (nc components):
// forward out[..k..] = sum_i Q[i][k] in[..i..]; inverse out[..i..] =
// sum_k Q[i][k] in[..k..].  With scale, each output is divided by its
// eigenvalue sum A a + sum_d lam_d (zero where that sum is below the
// component's tolerance).  Q of component c, direction axis, at
// Q + c * qs + qo[axis] (row-major n x n); eigenvalues of (c, d) at
// lam + c * ls + lo[d], the tolerance at lam + c * ls + ls - 1.
template <class T>
struct SepArgs
{
    int n[3];
    int dim;
    int axis;
    int inverse;
    int scale;
    const T* in;
    T* out;
    const T* Q;
    long long qs;
    int qo[3];
    const double* lam;
    long long ls;
    int lo[3];
    double aa;
    long long N;
    unsigned act;
};

__device__ __forceinline__ double sep_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float sep_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <class T>
__global__ void __launch_bounds__(256) mg_sep_transform_kernel(SepArgs<T> A)
{
    const int c = blockIdx.y;
    if (!((A.act >> c) & 1u)) return;
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= A.N) return;
    const int n0 = A.n[0], n1 = A.n[1];
    const long long r = q / n0;
    const int i = (int)(q - r * n0);
    const int k = (int)(r / n1);
    const int j = (int)(r - (long long)k * n1);
    const int ax = A.axis, n = A.n[ax];
    const int o = ax == 0 ? i : (ax == 1 ? j : k);
    const long long stride = ax == 0 ? 1 : (ax == 1 ? (long long)n0 : (long long)n0 * n1);
    const T* __restrict__ in = A.in + (long long)c * A.N + q - (long long)o * stride;
    const T* __restrict__ Qc = A.Q + (long long)c * A.qs + A.qo[ax];
    T acc = T(0);
    if (!A.inverse)
        for (int m = 0; m < n; ++m) acc = sep_fma(Qc[m * n + o], in[(long long)m * stride], acc);
    else
        for (int m = 0; m < n; ++m) acc = sep_fma(Qc[o * n + m], in[(long long)m * stride], acc);
    if (A.scale)
    {
        const double* lc = A.lam + (long long)c * A.ls;
        const int id[3] = {i, j, k};
        double den = A.aa;
        for (int d = 0; d < A.dim; ++d) den += lc[A.lo[d] + id[d]];
        const double tol = lc[A.ls - 1];
        acc = (fabs(den) > tol) ? T((double)acc / den) : T(0);
    }
    A.out[(long long)c * A.N + q] = acc;
}

// LDS-tiled form of the same products for n <= kSepTileN.  A workgroup
// takes kSepTI lines of the direction (axis 0: lines contiguous in memory;
// axes 1, 2: kSepTI consecutive inner indices of one outer index, the lines
// then strided by S = product of the lower extents) and kSepTK of the n
// outputs per line; the lines' inputs and the kSepTK matrix columns
// C[m][k] (forward Q[m][k], inverse Q[k][m]) sit in LDS.  A thread owns line
// t % kSepTI and outputs k0 + t / kSepTI + 4 u.  Same sums in the same order
// (m ascending, one fma per term) as mg_sep_transform_kernel.  Axis 0 hands
// its outputs back through LDS so that the stores are contiguous.
constexpr int kSepTileN = 64;
constexpr int kSepTI = 64;
constexpr int kSepTK = 16;
constexpr int kSepKG = 256 / kSepTI;        // thread groups over k
constexpr int kSepPK = kSepTK / kSepKG;     // outputs per thread

template <class T>
__device__ __forceinline__ T sep_scale(const SepArgs<T>& A, int c, long long q, T acc)
{
    const int n0 = A.n[0], n1 = A.n[1];
    const long long r = q / n0;
    const int i = (int)(q - r * n0);
    const int k = (int)(r / n1);
    const int j = (int)(r - (long long)k * n1);
    const double* lc = A.lam + (long long)c * A.ls;
    const int id[3] = {i, j, k};
    double den = A.aa;
    for (int d = 0; d < A.dim; ++d) den += lc[A.lo[d] + id[d]];
    return (fabs(den) > lc[A.ls - 1]) ? T((double)acc / den) : T(0);
}

template <class T, bool AX0>
__global__ void __launch_bounds__(256) mg_sep_tile_kernel(SepArgs<T> A)
{
    constexpr int P = kSepTI + 1;                  // padded row of the tile
    __shared__ T Tl[kSepTileN * P];
    __shared__ T Cs[kSepTileN * kSepTK];
    const int c = blockIdx.y;
    if (!((A.act >> c) & 1u)) return;
    const int ax = A.axis, n = A.n[ax];
    const int nk = (n + kSepTK - 1) / kSepTK;
    const int kt = (int)(blockIdx.x % (unsigned)nk);
    const long long bt = (long long)(blockIdx.x / (unsigned)nk);
    const int k0 = kt * kSepTK, tk = min(kSepTK, n - k0);
    const T* __restrict__ Qc = A.Q + (long long)c * A.qs + A.qo[ax];
    for (int q = threadIdx.x; q < n * kSepTK; q += 256)
    {
        const int m = q / kSepTK, kk = q - m * kSepTK;
        if (kk < tk) Cs[q] = A.inverse ? Qc[(k0 + kk) * n + m] : Qc[m * n + k0 + kk];
    }
    // tile rows m, columns = the workgroup's lines
    long long S, base;
    int ti;
    const T* __restrict__ in = A.in + (long long)c * A.N;
    if constexpr (AX0)
    {
        S = 1;
        const long long nlines = A.N / n, l0 = bt * kSepTI;
        ti = (int)min((long long)kSepTI, nlines - l0);
        base = l0 * n;
        for (int q = threadIdx.x; q < ti * n; q += 256)
        {
            const int i = q / n, m = q - i * n;
            Tl[m * P + i] = in[base + q];
        }
    } else
    {
        S = ax == 1 ? (long long)A.n[0] : (long long)A.n[0] * A.n[1];
        const long long nti = (S + kSepTI - 1) / kSepTI;
        const long long o = bt / nti, i0 = (bt - o * nti) * kSepTI;
        ti = (int)min((long long)kSepTI, S - i0);
        base = o * n * S + i0;
        for (int q = threadIdx.x; q < n * kSepTI; q += 256)
        {
            const int m = q / kSepTI, i = q - m * kSepTI;
            if (i < ti) Tl[m * P + i] = in[base + (long long)m * S + i];
        }
    }
    __syncthreads();
    const int t = threadIdx.x, i = t % kSepTI, g = t / kSepTI;
    T acc[kSepPK];
#pragma unroll
    for (int u = 0; u < kSepPK; ++u) acc[u] = T(0);
    if (i < ti)
        for (int m = 0; m < n; ++m)
        {
            const T x = Tl[m * P + i];
#pragma unroll
            for (int u = 0; u < kSepPK; ++u) acc[u] = sep_fma(Cs[m * kSepTK + g + kSepKG * u], x, acc[u]);
        }
    T* __restrict__ out = A.out + (long long)c * A.N;
    if constexpr (AX0)
    {
        // outputs (line i, k) through LDS: contiguous stores along k
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kSepPK; ++u)
        {
            const int kk = g + kSepKG * u;
            if (i < ti && kk < tk) Tl[kk * P + i] = acc[u];
        }
        __syncthreads();
        for (int q = threadIdx.x; q < ti * tk; q += 256)
        {
            const int li = q / tk, kk = q - li * tk;
            const long long cell = base + (long long)li * n + k0 + kk;
            const T v = Tl[kk * P + li];
            out[cell] = A.scale ? sep_scale(A, c, cell, v) : v;
        }
    } else
    {
        if (i >= ti) return;
#pragma unroll
        for (int u = 0; u < kSepPK; ++u)
        {
            const int kk = g + kSepKG * u;
            if (kk >= tk) continue;
            const long long cell = base + (long long)(k0 + kk) * S + i;
            out[cell] = A.scale ? sep_scale(A, c, cell, acc[u]) : acc[u];
        }
    }
}

// Division of the transformed right-hand side by the eigenvalue sums, for
// the products that run as library GEMMs (the tiled kernels fold it into
// their last forward product): same expression as sep_scale.
template <class T>
__global__ void __launch_bounds__(256) mg_sep_scale_kernel(SepArgs<T> A, int c, T* x)
{
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= A.N) return;
    x[q] = sep_scale(A, c, q, x[q]);
}

// Column-major GEMMs (rocBLAS) in the precision T.
inline void sep_gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, int k,
                     const double* A, int lda, long long sa, const double* B, int ldb, long long sb, double* C,
                     int ldc, long long sc, int batch)
{
    const double one = 1.0, zero = 0.0;
    rocblas_status st = (batch == 1)
                            ? rocblas_dgemm(h, ta, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc)
                            : rocblas_dgemm_strided_batched(h, ta, tb, m, n, k, &one, A, lda, sa, B, ldb, sb, &zero,
                                                            C, ldc, sc, batch);
    if (st != rocblas_status_success) throw std::runtime_error("rocblas_dgemm failed (" + std::to_string((int)st) + ")");
}
inline void sep_gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n, int k,
                     const float* A, int lda, long long sa, const float* B, int ldb, long long sb, float* C, int ldc,
                     long long sc, int batch)
{
    const float one = 1.0f, zero = 0.0f;
    rocblas_status st = (batch == 1)
                            ? rocblas_sgemm(h, ta, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc)
                            : rocblas_sgemm_strided_batched(h, ta, tb, m, n, k, &one, A, lda, sa, B, ldb, sb, &zero,
                                                            C, ldc, sc, batch);
    if (st != rocblas_status_success) throw std::runtime_error("rocblas_sgemm failed (" + std::to_string((int)st) + ")");
}

// One product along axis d of a component's level array as a GEMM.  The
// array is x fastest: with S the product of the extents below d and O
// the product of those above, entry (s, m, o) sits at s + S (m + n o).
// Forward out(.., k, ..) = sum_m Q[m][k] in(.., m, ..), inverse
// out(.., i, ..) = sum_k Q[i][k] in(.., k, ..); Q is row-major n x n, which
// column-major reads as Q^T.  d = 0 (S = 1): out (n x O) = op(Q^T) in;
// d > 0: for each o, out_o (S x n) = in_o op(Q^T)^T, O batches.
template <class T>
void sep_gemm_axis(rocblas_handle h, const int* ext, int dim, int d, bool inverse, const T* Q, const T* in, T* out)
{
    const int n = ext[d];
    long long S = 1, O = 1;
    for (int q = 0; q < d; ++q) S *= ext[q];
    for (int q = d + 1; q < dim; ++q) O *= ext[q];
    if (d == 0)
        sep_gemm(h, inverse ? rocblas_operation_transpose : rocblas_operation_none, rocblas_operation_none, n, (int)O,
                 n, Q, n, 0, in, n, 0, out, n, 0, 1);
    else
        sep_gemm(h, rocblas_operation_none, inverse ? rocblas_operation_none : rocblas_operation_transpose, (int)S, n,
                 n, in, (int)S, S * n, Q, n, 0, out, (int)S, S * n, (int)O);
}

} // namespace vd
#endif
