// Plane-marching fused smoothing passes of the 3D constant-coefficient
// multigrid levels (gfx950).
//
// A V-cycle level does, per pass of nu = 2 red-black sweeps, four colour
// stages (red, black, red, black).  As separate launches every stage
// streams the whole correction through HBM (half of each cache line is
// written, the line goes back whole), and the residual/restriction and the
// prolongation are passes of their own.  Here one workgroup owns a tile of
// 64 x 16 cells in (x, y) and a range of z planes, and marches up z with the
// four stages one plane apart: at step L
//     pre  pass:  stage 1 on plane L (from x = 0), stage s on plane L-s+1,
//                 residual + restriction on plane L-4;
//     post pass:  load plane L (x + prolongated coarse correction),
//                 stage s on plane L-s, write plane L-4.
// Stage s only needs stage s-1 values of the planes next to it, which the
// earlier stages of the same step or earlier steps produced.  The tile is
// grown by a 4-cell halo in x and y, recomputed redundantly: a stage
// updates the tile grown by (4 - s) (pre: 5 - s), which is exactly what the
// next stage reads, so every interior value equals the separate launches'
// value bit for bit (same operations in the same order: stencil_core,
// prolong_value, and the residual sum order of mg_resid_restrict_kernel).
// x planes live in a 6-plane LDS ring, rhs planes in a 5-plane ring, the
// pre pass's tile residuals in one more plane (160 KB in FP64: one
// workgroup per CU, 14 waves).
//
// Memory per pass per component, in units of the level's array: pre reads
// f once (plus halo) and writes x and f/8; post reads x, f and x_coarse/8
// and writes x once — against about 2.6 units per colour stage, plus the
// residual/restriction and prolongation passes, of the separate kernels.
#ifndef VD_MG_FUSED_H_
#define VD_MG_FUSED_H_

namespace vd {

constexpr int kFTX = 64, kFTY = 16, kFH = 4;                    // tile, halo
constexpr int kFRX = kFTX + 2 * kFH, kFRY = kFTY + 2 * kFH;     // region 72 x 24
constexpr int kFPS = kFRX * kFRY;                               // region plane
constexpr int kFNX = 6, kFNF = 5;                               // ring planes (x, f)
constexpr int kFPairs = (kFRX / 2) * kFRY;                      // x-pairs per plane
constexpr int kFThreads = 896;                                  // 14 waves >= kFPairs
static_assert(kFPairs <= kFThreads, "one x-pair per thread");
static_assert((kFTX / 2) * (kFTY / 2) <= kFThreads, "one coarse cell per thread");
static_assert(kFH % 2 == 0 && kFTX % 2 == 0 && kFTY % 2 == 0, "pairs and children aligned to even x");

template <class T>
struct FusedArgs
{
    MGView<T> v;                 // this level
    const LevelBC* bcs;          // this level's boundary model
    int pm;                      // periodic mask
    const T* f;                  // rhs, nc components
    const T* xin;                // post: pre-smoothed correction
    MGView<T> vc;                // post: coarse level
    const T* xcg;                // post: coarse correction with one ghost layer
                                 // ((c0+2) x (c1+2) x (c2+2) per component)
    T* xout;                     // correction after the pass
    T* fc;                       // pre: coarse rhs (restricted residual)
    long long Nc;                // pre: coarse cells per component
    int cn0, cn1;                // pre: coarse extents in x, y
    int zc, nchunk;              // z planes per workgroup, chunks per component
    unsigned active;
    // post on the finest level: add the result to the solution instead of
    // writing xout (component c at sol + compmap[c] * sstride; null: xout)
    double* sol;
    long long sstride;
    const int* compmap;
};

__device__ __forceinline__ int fz_slot(int p, int np)
{
    const int r = p % np;
    return r < 0 ? r + np : r;
}

// Per-workgroup geometry shared by both passes.
struct FusedTile
{
    int n0, n1, n2;
    bool px, py, pz;
    int x0, y0, ix1, iy1, z0, z1;   // interior box (x, y clipped) and z range
    int rx0, ry0;                   // region origin (global, unwrapped)
    __device__ bool alive_x(int gx) const { return px || (gx >= 0 && gx < n0); }
    __device__ bool alive_y(int gy) const { return py || (gy >= 0 && gy < n1); }
    __device__ bool alive_z(int gz) const { return pz || (gz >= 0 && gz < n2); }
    // cells a pass loads: in the domain (or a periodic image) and within the
    // clipped interior grown by the halo, the most any stage reads; this also
    // keeps a periodic image within one period of the domain (n >= kFH)
    __device__ bool load_x(int gx) const { return alive_x(gx) && gx >= x0 - kFH && gx < ix1 + kFH; }
    __device__ bool load_y(int gy) const { return alive_y(gy) && gy >= y0 - kFH && gy < iy1 + kFH; }
    __device__ static int wrap(int g, int n) { return g < 0 ? g + n : (g >= n ? g - n : g); }
    __device__ long long cell(int gx, int gy, int gz) const
    {
        return (long long)wrap(gx, n0) + (long long)n0 * ((long long)wrap(gy, n1) + (long long)n1 * wrap(gz, n2));
    }
};

template <class T>
__device__ __forceinline__ FusedTile fused_tile(const FusedArgs<T>& A, const dim3& b)
{
    FusedTile t;
    t.n0 = A.v.n[0];
    t.n1 = A.v.n[1];
    t.n2 = A.v.n[2];
    t.px = A.pm & 1;
    t.py = (A.pm >> 1) & 1;
    t.pz = (A.pm >> 2) & 1;
    t.x0 = (int)b.x * kFTX;
    t.y0 = (int)b.y * kFTY;
    t.ix1 = min(t.x0 + kFTX, t.n0);
    t.iy1 = min(t.y0 + kFTY, t.n1);
    const int chunk = (int)b.z % A.nchunk;
    t.z0 = chunk * A.zc;
    t.z1 = min(t.z0 + A.zc, t.n2);
    t.rx0 = t.x0 - kFH;
    t.ry0 = t.y0 - kFH;
    return t;
}

// Operator row of the cell at region (ra, rb) on plane q, reading the x
// ring.  (A variant that skipped the side models for cells away from the
// domain sides measured slower: waves holding both kinds ran both paths.)
template <class T>
__device__ __forceinline__ void fused_row(const MGView<T>& v, const SideBC (*bc)[2], const FusedTile& t,
                                          const T* ring, int ra, int rb, int q, T bcoef, T& xc, T& Lx, T& dg)
{
    const int gx = t.rx0 + ra, gy = t.ry0 + rb;
    const T* P = ring + fz_slot(q, kFNX) * kFPS;
    const T* Pm = ring + fz_slot(q - 1, kFNX) * kFPS;
    const T* Pp = ring + fz_slot(q + 1, kFNX) * kFPS;
    const int id = rb * kFRX + ra;
    // every read is inside the ring (stage cells sit at least one cell in
    // from the region edge); values at a domain side are replaced below
    xc = P[id];
    T xmr[3], xpr[3];
    xmr[0] = P[id - 1];
    xpr[0] = P[id + 1];
    xmr[1] = P[id - kFRX];
    xpr[1] = P[id + kFRX];
    xmr[2] = Pm[id];
    xpr[2] = Pp[id];
    const T bm[3] = {bcoef, bcoef, bcoef};
    bool lob[3], hib[3];
    lob[0] = !t.px && gx == 0;
    hib[0] = !t.px && gx == t.n0 - 1;
    lob[1] = !t.py && gy == 0;
    hib[1] = !t.py && gy == t.n1 - 1;
    lob[2] = !t.pz && q == 0;
    hib[2] = !t.pz && q == t.n2 - 1;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        if (lob[d]) xmr[d] = xc;
        if (hib[d]) xpr[d] = xc;
    }
    stencil_core<3, T>(v, bc, xc, xmr, xpr, lob, hib, bm, bm, Lx, dg);
}

// Diagonal of a row with no domain side (the same arithmetic as stencil_core).
template <class T>
__device__ __forceinline__ T fused_inner_diag(const MGView<T>& v, const SideBC (*bc)[2], T bcoef)
{
    const bool no[3] = {false, false, false};
    const T z[3] = {T(0), T(0), T(0)}, bm[3] = {bcoef, bcoef, bcoef};
    T Lx, dg;
    stencil_core<3, T>(v, bc, T(0), z, z, no, no, bm, bm, Lx, dg);
    return dg;
}

// Colour stage on plane q: the thread's x-pair updates its cell of colour
// col when it lies in the interior grown by g (and in the domain).
template <class T>
__device__ __forceinline__ void fused_stage(const MGView<T>& v, const SideBC (*bc)[2], const FusedTile& t, T* ring,
                                            const T* fring, int ra0, int rb, int q, int col, int g, int zg, T bcoef)
{
    if (q < t.z0 - zg || q >= t.z1 + zg || !t.alive_z(q)) return;
    const int gy = t.ry0 + rb;
    if (gy < t.y0 - g || gy >= t.iy1 + g || !t.alive_y(gy)) return;
    const int ra = ra0 + ((gy + q + col) & 1);   // rx0 is even
    const int gx = t.rx0 + ra;
    if (gx < t.x0 - g || gx >= t.ix1 + g || !t.alive_x(gx)) return;
    T xc, Lx, dg;
    fused_row<T>(v, bc, t, ring, ra, rb, q, bcoef, xc, Lx, dg);
    const T fv = fring[fz_slot(q, kFNF) * kFPS + rb * kFRX + ra];
    ring[fz_slot(q, kFNX) * kFPS + rb * kFRX + ra] = xc + (fv - Lx) / dg;
}

// Pre-smoothing pass from x = 0 (nu1 = 2) fused with residual + restriction:
// writes the correction (interior) and the coarse rhs.  The residuals of
// the tile's cells go to an LDS plane; the thread owning a coarse cell sums
// its children there one step later (two planes per coarse cell).
template <class T>
__global__ void __launch_bounds__(kFThreads) mg_fused_pre_kernel(FusedArgs<T> A)
{
    __shared__ T ring[kFNX * kFPS];
    __shared__ T fring[kFNF * kFPS];
    __shared__ T rplane[kFTX * kFTY];
    const dim3 b = xcd_block();
    const int c = __builtin_amdgcn_readfirstlane((int)b.z / A.nchunk);
    if (!((A.active >> c) & 1u)) return;
    const FusedTile t = fused_tile(A, b);
    const MGView<T>& v = A.v;
    // the side models of this component, read once (uniform)
    SideBC bc[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) bc[d][s2] = A.bcs->s[c][d][s2];
    const T bcoef = T(v.bcomp ? v.bcomp[c] : v.bconst);
    const T dgi = fused_inner_diag<T>(v, bc, bcoef);
    const long long cb = (long long)c * v.Nl;
    const T* __restrict__ f = A.f + cb;
    const int tid = threadIdx.x;
    const bool pair = tid < kFPairs;
    const int ra0 = 2 * (tid % (kFRX / 2)), rb = tid / (kFRX / 2);
    const int gy = t.ry0 + rb;
    // coarse cell owned by this thread (restriction)
    const int ci = tid % (kFTX / 2), cj = tid / (kFTX / 2);
    const bool owner = tid < (kFTX / 2) * (kFTY / 2) && (t.x0 + 2 * ci < t.ix1) && (t.y0 + 2 * cj < t.iy1);
    T rsum = T(0);
    // sum of the children of plane q (after the barrier that follows their
    // residuals), in the order of mg_resid_restrict_kernel
    auto restrict_plane = [&](int q) __attribute__((always_inline)) {
        if (!owner || q < t.z0 || q >= t.z1) return;
        if (((q - t.z0) & 1) == 0) rsum = T(0);
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) rsum += rplane[(2 * cj + jj) * kFTX + 2 * ci + ii];
        if ((q - t.z0) & 1)
        {
            const int I = t.x0 / 2 + ci, J = t.y0 / 2 + cj, K = q >> 1;
            A.fc[(long long)c * A.Nc + (long long)I + (long long)A.cn0 * ((long long)J + (long long)A.cn1 * K)] =
                rsum / T(8);
        }
    };
    // prefetched rhs pair of the next plane
    T fn0 = T(0), fn1 = T(0);
    auto fetch_f = [&](int p) __attribute__((always_inline)) {
        fn0 = fn1 = T(0);
        if (!pair || !t.alive_z(p) || !t.load_y(gy)) return;
        const int gx = t.rx0 + ra0;
        if (t.load_x(gx)) fn0 = f[t.cell(gx, gy, p)];
        if (t.load_x(gx + 1)) fn1 = f[t.cell(gx + 1, gy, p)];
    };
    const int Lb = t.z0 - kFH, Le = t.z1 + kFH;   // steps L in [Lb, Le)
    fetch_f(Lb);
    for (int L = Lb; L < Le; ++L)
    {
        // stage 1 on plane L from x = 0: colour 0 gets f/diag, colour 1 zero
        if (pair)
        {
            T* Fp = fring + fz_slot(L, kFNF) * kFPS + rb * kFRX + ra0;
            Fp[0] = fn0;
            Fp[1] = fn1;
            T* Xp = ring + fz_slot(L, kFNX) * kFPS + rb * kFRX + ra0;
            T x0v = T(0), x1v = T(0);
            if (t.alive_z(L) && t.load_y(gy))
            {
                const int off = (gy + L) & 1;   // colour-0 cell of the pair
                const int gx = t.rx0 + ra0 + off;
                if (t.load_x(gx))
                {
                    bool lob[3], hib[3];
                    lob[0] = !t.px && gx == 0;
                    hib[0] = !t.px && gx == t.n0 - 1;
                    lob[1] = !t.py && gy == 0;
                    hib[1] = !t.py && gy == t.n1 - 1;
                    lob[2] = !t.pz && L == 0;
                    hib[2] = !t.pz && L == t.n2 - 1;
                    T dg = dgi;
                    if (lob[0] || hib[0] || lob[1] || hib[1] || lob[2] || hib[2])
                    {
                        const T zr[3] = {T(0), T(0), T(0)};
                        const T bm[3] = {bcoef, bcoef, bcoef};
                        T Lx;
                        stencil_core<3, T>(v, bc, T(0), zr, zr, lob, hib, bm, bm, Lx, dg);
                    }
                    const T val = T(0) + ((off ? fn1 : fn0) - T(0)) / dg;
                    if (off) x1v = val;
                    else x0v = val;
                }
            }
            Xp[0] = x0v;
            Xp[1] = x1v;
        }
        restrict_plane(L - 5);
        if (L + 1 < Le) fetch_f(L + 1);
        __syncthreads();
        if (pair) fused_stage<T>(v, bc, t, ring, fring, ra0, rb, L - 1, 1, 3, 3, bcoef);
        __syncthreads();
        if (pair) fused_stage<T>(v, bc, t, ring, fring, ra0, rb, L - 2, 0, 2, 2, bcoef);
        __syncthreads();
        if (pair) fused_stage<T>(v, bc, t, ring, fring, ra0, rb, L - 3, 1, 1, 1, bcoef);
        __syncthreads();
        // plane L-4 is final: write it, residuals of the tile's cells
        const int q = L - 4;
        if (q >= t.z0 && q < t.z1)
        {
            if (pair && gy >= t.y0 && gy < t.iy1)
            {
                const int gx = t.rx0 + ra0;
                if (gx >= t.x0 && gx < t.ix1)
                {
                    const T* Xp = ring + fz_slot(q, kFNX) * kFPS + rb * kFRX + ra0;
                    const long long o = cb + (long long)gx + (long long)t.n0 * ((long long)gy + (long long)t.n1 * q);
                    A.xout[o] = Xp[0];
                    A.xout[o + 1] = Xp[1];
                }
            }
            for (int e = tid; e < kFTX * kFTY; e += kFThreads)
            {
                const int ex = e % kFTX, ey = e / kFTX;
                if (t.x0 + ex >= t.ix1 || t.y0 + ey >= t.iy1) continue;
                const int ra = kFH + ex, rbb = kFH + ey;
                T xc, Lx, dg;
                fused_row<T>(v, bc, t, ring, ra, rbb, q, bcoef, xc, Lx, dg);
                rplane[e] = fring[fz_slot(q, kFNF) * kFPS + rbb * kFRX + ra] - Lx;
            }
        }
        __syncthreads();
    }
    restrict_plane(Le - 5);
}

// Post-smoothing pass (nu2 = 2) fused with the prolongation of the coarse
// correction: x_out = smooth(x_in + P x_coarse) on the interior.
template <class T>
__global__ void __launch_bounds__(kFThreads) mg_fused_post_kernel(FusedArgs<T> A)
{
    __shared__ T ring[kFNX * kFPS];
    __shared__ T fring[kFNF * kFPS];
    const dim3 b = xcd_block();
    const int c = __builtin_amdgcn_readfirstlane((int)b.z / A.nchunk);
    if (!((A.active >> c) & 1u)) return;
    const FusedTile t = fused_tile(A, b);
    const MGView<T>& v = A.v;
    // the side models of this component, read once (uniform)
    SideBC bc[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) bc[d][s2] = A.bcs->s[c][d][s2];
    const T bcoef = T(v.bcomp ? v.bcomp[c] : v.bconst);
    const long long cb = (long long)c * v.Nl;
    const T* __restrict__ f = A.f + cb;
    const T* __restrict__ xin = A.xin + cb;
    const int g0 = A.vc.n[0] + 2, g1 = A.vc.n[1] + 2;
    const T* __restrict__ xcb = A.xcg + (long long)c * g0 * g1 * (A.vc.n[2] + 2);
    const int tid = threadIdx.x;
    const bool pair = tid < kFPairs;
    const int ra0 = 2 * (tid % (kFRX / 2)), rb = tid / (kFRX / 2);
    const int gy = t.ry0 + rb;
    // prefetched raw values of the next plane: the pair's x and f, and the
    // coarse neighbourhood of its parent (I-1..I+1 in x, J and J+sy, K and
    // K+sz: the pair is the two x-children of one parent, rx0 being even)
    T xn0 = T(0), xn1 = T(0), fn0 = T(0), fn1 = T(0);
    T nb[2][2][3];
    bool live0 = false, live1 = false;
    auto fetch = [&](int p) __attribute__((always_inline)) {
        live0 = live1 = false;
        if (!pair || !t.alive_z(p) || !t.load_y(gy)) return;
        const int gx = t.rx0 + ra0;
        live0 = t.load_x(gx);
        live1 = t.load_x(gx + 1);
        if (!live0 && !live1) return;
        const int i = FusedTile::wrap(live0 ? gx : gx + 1, t.n0) & ~1;
        const int j = FusedTile::wrap(gy, t.n1), k = FusedTile::wrap(p, t.n2);
        const long long o = (long long)i + (long long)t.n0 * ((long long)j + (long long)t.n1 * k);
        if (live0)
        {
            xn0 = xin[o];
            fn0 = f[o];
        }
        if (live1)
        {
            xn1 = xin[o + 1];
            fn1 = f[o + 1];
        }
        const int I = i >> 1, J = j >> 1, K = k >> 1;
        const int sy = (j & 1) ? 1 : -1, sz = (k & 1) ? 1 : -1;
        // ghosted coarse index of (I-1, J, K)
        const long long co = (long long)I + (long long)g0 * ((long long)(J + 1) + (long long)g1 * (K + 1));
#pragma unroll
        for (int cc = 0; cc < 2; ++cc)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    nb[cc][bb][a] = xcb[co + a + (long long)g0 * (bb * sy + (long long)g1 * cc * sz)];
    };
    // child ii of the parent: the weighted sum of mg_prolong_kernel, in its
    // order (cc, bb, aa), neighbour I + aa*sx with sx = -1 (ii = 0) or +1
    auto child = [&](int ii) __attribute__((always_inline)) {
        const int sx = ii ? 1 : -1;
        T acc = T(0);
#pragma unroll
        for (int cc = 0; cc < 2; ++cc)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int aa = 0; aa < 2; ++aa)
                {
                    T w = aa ? T(0.25) : T(0.75);
                    w *= bb ? T(0.25) : T(0.75);
                    w *= cc ? T(0.25) : T(0.75);
                    acc += w * nb[cc][bb][1 + aa * sx];
                }
        return acc;
    };
    const int Lb = t.z0 - kFH, Le = t.z1 + kFH;
    fetch(Lb);
    for (int L = Lb; L < Le; ++L)
    {
        if (pair)
        {
            T* Xp = ring + fz_slot(L, kFNX) * kFPS + rb * kFRX + ra0;
            T* Fp = fring + fz_slot(L, kFNF) * kFPS + rb * kFRX + ra0;
            Xp[0] = live0 ? xn0 + child(0) : T(0);
            Xp[1] = live1 ? xn1 + child(1) : T(0);
            Fp[0] = live0 ? fn0 : T(0);
            Fp[1] = live1 ? fn1 : T(0);
        }
        if (L + 1 < Le) fetch(L + 1);
        __syncthreads();
        if (pair) fused_stage<T>(v, bc, t, ring, fring, ra0, rb, L - 1, 0, 3, 3, bcoef);
        __syncthreads();
        if (pair) fused_stage<T>(v, bc, t, ring, fring, ra0, rb, L - 2, 1, 2, 2, bcoef);
        __syncthreads();
        if (pair) fused_stage<T>(v, bc, t, ring, fring, ra0, rb, L - 3, 0, 1, 1, bcoef);
        __syncthreads();
        if (pair)
        {
            fused_stage<T>(v, bc, t, ring, fring, ra0, rb, L - 4, 1, 0, 0, bcoef);
            const int q = L - 4;
            // the pair's two cells were last written by this thread
            if (q >= t.z0 && q < t.z1 && gy >= t.y0 && gy < t.iy1)
            {
                const int gx = t.rx0 + ra0;
                if (gx >= t.x0 && gx < t.ix1)
                {
                    const T* Xp = ring + fz_slot(q, kFNX) * kFPS + rb * kFRX + ra0;
                    const long long oc = (long long)gx + (long long)t.n0 * ((long long)gy + (long long)t.n1 * q);
                    if (A.sol)
                    {
                        // mg_add_kernel's update, x = x + e
                        double* sp = A.sol + (long long)A.compmap[c] * A.sstride + oc;
                        sp[0] = sp[0] + double(Xp[0]);
                        sp[1] = sp[1] + double(Xp[1]);
                    } else
                    {
                        A.xout[cb + oc] = Xp[0];
                        A.xout[cb + oc + 1] = Xp[1];
                    }
                }
            }
        }
        __syncthreads();
    }
}

// Coarse correction with one ghost layer for the post pass: every entry is
// cval_x of its (possibly out-of-range) index, the value mg_prolong_kernel
// and prolong_value read there.
template <class T>
__global__ void __launch_bounds__(256) mg_coarse_ghost_kernel(MGView<T> vc, const LevelBC* __restrict__ bcs, int pm,
                                                              const T* __restrict__ xc, T* __restrict__ xg,
                                                              unsigned active)
{
    const int g0 = vc.n[0] + 2, g1 = vc.n[1] + 2, g2 = vc.n[2] + 2;
    const int c = (int)blockIdx.z / g2;
    const int K = (int)blockIdx.z - c * g2;
    if (!((active >> c) & 1u)) return;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= g0 || J >= g1) return;
    const long long o = (long long)c * g0 * g1 * g2 + (long long)I + (long long)g0 * ((long long)J + (long long)g1 * K);
    const T* xcb = xc + (long long)c * vc.Nl;
    if (I >= 1 && I <= vc.n[0] && J >= 1 && J <= vc.n[1] && K >= 1 && K <= vc.n[2])
        xg[o] = xcb[(long long)(I - 1) + (long long)vc.n[0] * ((long long)(J - 1) + (long long)vc.n[1] * (K - 1))];
    else
        xg[o] = cval_x<3, T>(vc, bcs->s[c], pm, xcb, I - 1, J - 1, K - 1);
}

} // namespace vd
#endif
