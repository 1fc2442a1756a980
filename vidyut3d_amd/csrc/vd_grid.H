// Grid geometry, boundary descriptors and launch helpers shared by all
// vidyut3d_amd kernels (gfx950).
//
// State layout: the vidyut3d MultiFab layout of one box without ghost cells,
// offset (i + nx*(j + ny*k)) + comp*N, x fastest, component outermost.
// Ghost values are never stored: a kernel that needs the value beyond the
// domain maps the index the way the reference's FillPatch fills ghosts
// (Source/Vidyut.cpp:118-130): wrap for periodic directions, nearest valid
// cell (foextrap) otherwise.
#ifndef VD_GRID_H_
#define VD_GRID_H_

#include <hip/hip_runtime.h>
#include <cstdint>

namespace vd {

constexpr int kMaxComp = 32;

struct Grid
{
    int n[3];        // cells per direction (1 for unused directions)
    int dim;         // 2 or 3
    double dx[3];
    double problo[3], probhi[3];
    int periodic[3];
    long long N;     // cells
    int off[3];      // global index of cell (0,0,0) on its level (0 on level 0)
    int coord;       // 0 Cartesian, 1 axisymmetric RZ (2D, x = r): rz_factors
    __host__ __device__ long long idx(int i, int j, int k) const
    {
        return (long long)i + (long long)n[0] * ((long long)j + (long long)n[1] * k);
    }
};

// Axisymmetric (RZ, coord_sys = 1, 2D: x = r) metric of the radial
// direction.  AMReX multiplies a by r_c, the radial face coefficients by
// r_face, the axial ones and the rhs by r_c; here each operator row is
// divided by its r_c instead (the same solution): the radial faces of cell
// i carry rm = r_{i-1/2} / r_c and rp = r_{i+1/2} / r_c, nothing else
// changes.  Every operator evaluation and boundary fold takes them from here.
__host__ __device__ __forceinline__ void rz_factors(double rlo, double dx, int i, double& rm, double& rp)
{
    const double rc = rlo + (i + 0.5) * dx;
    rm = (rc - 0.5 * dx) / rc;
    rp = (rc + 0.5 * dx) / rc;
}

// FillPatch ghost image of index q in a direction with extent n.
__host__ __device__ inline int ghost_map(int q, int n, int periodic)
{
    if (q >= 0 && q < n) return q;
    if (periodic)
    {
        int r = q % n;
        return r < 0 ? r + n : r;
    }
    return q < 0 ? 0 : n - 1;
}

// Linear-solver boundary kinds (LinOpBCType mapping of
// Source/PotentialSolve.cpp:95-157 and Source/ScalarSolve.cpp:467-529).
enum BcKind : int { BK_PER = 0, BK_DIR = 1, BK_NEU = 2, BK_INHNEU = 3, BK_ROBIN = 4, BK_CF = 5 };
// BK_CF: a fine level's coarse/fine side (ghost = 2/3 x0 - 1/5 x1 plus the
// parent part folded into the right-hand side).

// Homogeneous ghost model of one domain side for one component:
//   ghost = dep0 * x0 + dep1 * x1
// where x0 is the boundary cell and x1 the next interior cell.
struct SideBC
{
    int kind;
    double dep0, dep1;
};

// Per level: [comp][dir][side]
struct LevelBC
{
    SideBC s[kMaxComp][3][2];
};

// XCD-aware workgroup order.  gfx950 hands consecutive workgroups of a
// launch to its 8 XCDs in turn (workgroup p runs on XCD p % 8), and each XCD
// has its own L2.  Stencil kernels want neighbouring tiles (which share halo
// rows and planes) in the same L2, so the logical workgroup index gives XCD x
// one contiguous range of tiles: physical id p -> logical
//   (p % 8) * (B / 8) + min(p % 8, B % 8) + p / 8,
// a bijection of [0, B).  Kernels read their tile from xcd_block() instead
// of blockIdx.
__device__ __forceinline__ dim3 xcd_block()
{
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const unsigned B = gx * gy * gz;
    const unsigned p = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned q = B >> 3, r = B & 7u, x = p & 7u, t = p >> 3;
    const unsigned L = x * q + (x < r ? x : r) + t;
    dim3 b;
    b.x = L % gx;
    b.y = (L / gx) % gy;
    b.z = L / (gx * gy);
    return b;
}

// Launch shape for cell kernels: x fastest, one thread per cell.
struct Launch3
{
    dim3 grid, block;
};
inline Launch3 cell_launch(const int n[3], int bx = 64, int by = 4)
{
    Launch3 L;
    L.block = dim3(bx, by, 1);
    L.grid = dim3((n[0] + bx - 1) / bx, (n[1] + by - 1) / by, n[2]);
    return L;
}

} // namespace vd
#endif
