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

} // namespace vd
#endif
