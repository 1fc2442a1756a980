// Minimal AMReX-API-compatible surface for vidyut3d case plugins.
//
// Case directories (Chemistry.H, UserFunctions.H, UserSources.H, Prob.H,
// ProbParm.H) are written against a small subset of AMReX: amrex::Real,
// GpuArray, Array4, IntVect, Box, GeometryData, Math::powi, ParmParse,
// Gpu::htod_memcpy and the AMREX_* function-qualifier macros.  This header
// provides exactly that subset, natively, so those headers compile unchanged
// for gfx950 (hipcc) and for the host (g++).  Nothing here launches work.
//
// Array4 carries an optional ghost policy: a view over a ghost-free state
// array can be read at indices outside the valid box, and the index is mapped
// to the cell the reference's FillPatch would have copied there (foextrap:
// nearest valid cell; periodic: wrapped image).  Views over ordinary arrays
// use policy 0 and index directly.
#ifndef VD_AMREX_COMPAT_H_
#define VD_AMREX_COMPAT_H_

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <string>
#include <vector>
#include <array>
#include <map>
#include <algorithm>
#include <iostream>
#include <sstream>
#include <limits>
#include <stdexcept>
#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VD_HD __host__ __device__
#define VD_D __device__
#ifndef AMREX_USE_GPU
#define AMREX_USE_GPU 1
#endif
#ifndef AMREX_USE_HIP
#define AMREX_USE_HIP 1
#endif
#define AMREX_GPU_CONSTANT __constant__
#else
#define VD_HD
#define VD_D
#define AMREX_GPU_CONSTANT
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define AMREX_DEVICE_COMPILE 1
#else
#define AMREX_DEVICE_COMPILE 0
#endif

#ifndef AMREX_SPACEDIM
#error "AMREX_SPACEDIM must be defined (2 or 3) when compiling a case"
#endif

#define AMREX_GPU_HOST_DEVICE VD_HD
#define AMREX_GPU_DEVICE VD_D
#define AMREX_GPU_HOST
#define AMREX_FORCE_INLINE inline __attribute__((always_inline))
#define AMREX_INLINE inline
#define AMREX_RESTRICT __restrict__
#define AMREX_PRAGMA_SIMD
#define AMREX_GPU_DEVICE_MANAGED
#define AMREX_ALWAYS_ASSERT(x) ((void)0)
#define AMREX_ASSERT(x) ((void)0)
#define BL_PROFILE(x) ((void)0)

#if AMREX_SPACEDIM == 1
#define AMREX_D_DECL(a, b, c) a
#define AMREX_D_TERM(a, b, c) a
#elif AMREX_SPACEDIM == 2
#define AMREX_D_DECL(a, b, c) a, b
#define AMREX_D_TERM(a, b, c) a b
#else
#define AMREX_D_DECL(a, b, c) a, b, c
#define AMREX_D_TERM(a, b, c) a b c
#endif

namespace amrex {

using Real = double;
using Long = long long;

namespace literals {
constexpr Real operator""_rt(long double x) { return Real(x); }
constexpr Real operator""_rt(unsigned long long x) { return Real(x); }
} // namespace literals

template <class T>
using Vector = std::vector<T>;

template <class T, int N>
struct GpuArray
{
    T arr[N];
    VD_HD T& operator[](int i) { return arr[i]; }
    VD_HD const T& operator[](int i) const { return arr[i]; }
    VD_HD T* data() { return arr; }
    VD_HD const T* data() const { return arr; }
    VD_HD static constexpr int size() { return N; }
};

struct Dim3
{
    int x, y, z;
};

struct IntVect
{
    int vect[AMREX_SPACEDIM];
    VD_HD IntVect()
    {
        for (int d = 0; d < AMREX_SPACEDIM; ++d) vect[d] = 0;
    }
    VD_HD explicit IntVect(AMREX_D_DECL(int i, int j, int k))
    {
        vect[0] = i;
#if AMREX_SPACEDIM > 1
        vect[1] = j;
#endif
#if AMREX_SPACEDIM > 2
        vect[2] = k;
#endif
    }
    VD_HD int& operator[](int d) { return vect[d]; }
    VD_HD const int& operator[](int d) const { return vect[d]; }
    VD_HD static IntVect TheDimensionVector(int d)
    {
        IntVect v;
        v[d] = 1;
        return v;
    }
};

struct Box
{
    IntVect smallend, bigend;
    VD_HD Box() = default;
    VD_HD Box(IntVect const& lo, IntVect const& hi) : smallend(lo), bigend(hi) {}
    VD_HD const IntVect& smallEnd() const { return smallend; }
    VD_HD const IntVect& bigEnd() const { return bigend; }
    VD_HD int smallEnd(int d) const { return smallend[d]; }
    VD_HD int bigEnd(int d) const { return bigend[d]; }
    VD_HD const int* loVect() const { return smallend.vect; }
    VD_HD const int* hiVect() const { return bigend.vect; }
    VD_HD int length(int d) const { return bigend[d] - smallend[d] + 1; }
};

VD_HD inline Dim3 lbound(Box const& b)
{
    return Dim3{
        b.smallend[0],
#if AMREX_SPACEDIM > 1
        b.smallend[1],
#else
        0,
#endif
#if AMREX_SPACEDIM > 2
        b.smallend[2]
#else
        0
#endif
    };
}
VD_HD inline Dim3 ubound(Box const& b)
{
    return Dim3{
        b.bigend[0],
#if AMREX_SPACEDIM > 1
        b.bigend[1],
#else
        0,
#endif
#if AMREX_SPACEDIM > 2
        b.bigend[2]
#else
        0
#endif
    };
}

// Ghost policy bits for Array4 views over ghost-free state arrays.
// bit d (0..2): direction d is periodic (wrap); otherwise foextrap (clamp).
// bit 3: policy active.
enum : int { VD_GHOST_ACTIVE = 8 };

template <class T>
struct Array4
{
    T* p = nullptr;
    long long jstride = 0, kstride = 0, nstride = 0;
    Dim3 begin{0, 0, 0};
    Dim3 end{1, 1, 1}; // exclusive
    int ncomp = 1;
    int ghost = 0;     // ghost policy bits (see above)

    VD_HD Array4() = default;
    VD_HD Array4(T* a_p, Dim3 const& a_begin, Dim3 const& a_end, int a_ncomp)
        : p(a_p), begin(a_begin), end(a_end), ncomp(a_ncomp)
    {
        jstride = (long long)(end.x - begin.x);
        kstride = jstride * (long long)(end.y - begin.y);
        nstride = kstride * (long long)(end.z - begin.z);
    }
    template <class U>
    VD_HD Array4(Array4<U> const& o)
        : p(o.p), jstride(o.jstride), kstride(o.kstride), nstride(o.nstride),
          begin(o.begin), end(o.end), ncomp(o.ncomp), ghost(o.ghost)
    {
    }

    VD_HD static int map_idx(int i, int lo, int hi_excl, bool periodic)
    {
        if (i >= lo && i < hi_excl) return i;
        int n = hi_excl - lo;
        if (periodic)
        {
            int r = (i - lo) % n;
            if (r < 0) r += n;
            return lo + r;
        }
        return (i < lo) ? lo : hi_excl - 1;
    }

    VD_HD long long offset(int i, int j, int k) const
    {
        if (ghost & VD_GHOST_ACTIVE)
        {
            i = map_idx(i, begin.x, end.x, ghost & 1);
            j = map_idx(j, begin.y, end.y, ghost & 2);
            k = map_idx(k, begin.z, end.z, ghost & 4);
        }
        return (long long)(i - begin.x) + (long long)(j - begin.y) * jstride +
               (long long)(k - begin.z) * kstride;
    }
    VD_HD T& operator()(int i, int j, int k) const { return p[offset(i, j, k)]; }
    VD_HD T& operator()(int i, int j, int k, int n) const
    {
        return p[offset(i, j, k) + (long long)n * nstride];
    }
    VD_HD T& operator()(IntVect const& iv) const
    {
#if AMREX_SPACEDIM == 1
        return (*this)(iv[0], 0, 0);
#elif AMREX_SPACEDIM == 2
        return (*this)(iv[0], iv[1], 0);
#else
        return (*this)(iv[0], iv[1], iv[2]);
#endif
    }
    VD_HD T& operator()(IntVect const& iv, int n) const
    {
#if AMREX_SPACEDIM == 1
        return (*this)(iv[0], 0, 0, n);
#elif AMREX_SPACEDIM == 2
        return (*this)(iv[0], iv[1], 0, n);
#else
        return (*this)(iv[0], iv[1], iv[2], n);
#endif
    }
    VD_HD int nComp() const { return ncomp; }
};

struct GeometryData
{
    Real prob_lo[AMREX_SPACEDIM];
    Real prob_hi[AMREX_SPACEDIM];
    Real dx[AMREX_SPACEDIM];
    Box domain;
    int coord = 0;
    int is_periodic[AMREX_SPACEDIM];
    VD_HD const Real* ProbLo() const { return prob_lo; }
    VD_HD const Real* ProbHi() const { return prob_hi; }
    VD_HD const Real* CellSize() const { return dx; }
    VD_HD Real ProbLo(int d) const { return prob_lo[d]; }
    VD_HD Real ProbHi(int d) const { return prob_hi[d]; }
    VD_HD Real CellSize(int d) const { return dx[d]; }
    VD_HD const Box& Domain() const { return domain; }
    VD_HD int Coord() const { return coord; }
    VD_HD int isPeriodic(int d) const { return is_periodic[d]; }
};

using RealArray = std::array<Real, AMREX_SPACEDIM>;

struct RealVect
{
    Real vect[AMREX_SPACEDIM];
    VD_HD RealVect(AMREX_D_DECL(Real a, Real b, Real c))
    {
        vect[0] = a;
#if AMREX_SPACEDIM > 1
        vect[1] = b;
#endif
#if AMREX_SPACEDIM > 2
        vect[2] = c;
#endif
    }
    VD_HD Real& operator[](int d) { return vect[d]; }
    VD_HD const Real& operator[](int d) const { return vect[d]; }
};

namespace Math {
template <int N, class T>
VD_HD AMREX_FORCE_INLINE T powi(T x)
{
    if constexpr (N == 0)
    {
        return T(1);
    } else if constexpr (N < 0)
    {
        return T(1) / powi<-N>(x);
    } else if constexpr (N == 1)
    {
        return x;
    } else if constexpr (N % 2 == 0)
    {
        T h = powi<N / 2>(x);
        return h * h;
    } else
    {
        return x * powi<N - 1>(x);
    }
}
template <class T>
VD_HD AMREX_FORCE_INLINE T abs(T x)
{
    return x < T(0) ? -x : x;
}
VD_HD AMREX_FORCE_INLINE Real floor(Real x) { return ::floor(x); }
VD_HD AMREX_FORCE_INLINE Real ceil(Real x) { return ::ceil(x); }
VD_HD AMREX_FORCE_INLINE Real exp(Real x) { return ::exp(x); }
VD_HD AMREX_FORCE_INLINE Real log(Real x) { return ::log(x); }
VD_HD AMREX_FORCE_INLINE Real sqrt(Real x) { return ::sqrt(x); }
VD_HD AMREX_FORCE_INLINE Real pow(Real x, Real y) { return ::pow(x, y); }
constexpr Real pi() { return 3.14159265358979323846264338327950288; }
} // namespace Math

template <class T>
VD_HD AMREX_FORCE_INLINE T min(T a, T b)
{
    return a < b ? a : b;
}
template <class T>
VD_HD AMREX_FORCE_INLINE T max(T a, T b)
{
    return a > b ? a : b;
}
template <class T, class... Ts>
VD_HD AMREX_FORCE_INLINE T min(T a, T b, Ts... rest)
{
    return min(min(a, b), rest...);
}
template <class T, class... Ts>
VD_HD AMREX_FORCE_INLINE T max(T a, T b, Ts... rest)
{
    return max(max(a, b), rest...);
}

// ---------------------------------------------------------------------------
// Host-side services used by amrex_probinit and case Chemistry.cpp files.

// Parameter table shared by all ParmParse objects of the process; the
// library fills it from the inputs text before amrex_probinit runs.
struct ParmTable
{
    std::map<std::string, std::vector<std::string>> entries;
    static ParmTable& get()
    {
        static ParmTable t;
        return t;
    }
};

class ParmParse
{
public:
    explicit ParmParse(std::string const& prefix = "") : m_prefix(prefix) {}

    bool contains(std::string const& name) const
    {
        return ParmTable::get().entries.count(full(name)) > 0;
    }
    int countval(std::string const& name) const
    {
        auto it = ParmTable::get().entries.find(full(name));
        return it == ParmTable::get().entries.end() ? 0 : int(it->second.size());
    }
    template <class T>
    int query(std::string const& name, T& v, int ival = 0) const
    {
        auto it = ParmTable::get().entries.find(full(name));
        if (it == ParmTable::get().entries.end() || int(it->second.size()) <= ival) return 0;
        parse(it->second[ival], v);
        return 1;
    }
    template <class T>
    void get(std::string const& name, T& v, int ival = 0) const
    {
        if (!query(name, v, ival))
        {
            throw std::runtime_error("ParmParse: required parameter missing: " + full(name));
        }
    }
    template <class T>
    int queryarr(std::string const& name, std::vector<T>& v, int start = 0, int n = -1) const
    {
        auto it = ParmTable::get().entries.find(full(name));
        if (it == ParmTable::get().entries.end()) return 0;
        int avail = int(it->second.size()) - start;
        if (n < 0) n = avail;
        if (n > avail) n = avail;
        if (n < 0) n = 0;
        if (int(v.size()) < start + n) v.resize(start + n);
        for (int i = 0; i < n; ++i) parse(it->second[start + i], v[start + i]);
        return 1;
    }
    template <class T>
    void getarr(std::string const& name, std::vector<T>& v, int start = 0, int n = -1) const
    {
        if (!queryarr(name, v, start, n))
        {
            throw std::runtime_error("ParmParse: required array missing: " + full(name));
        }
    }
    template <class T, size_t N>
    int queryarr(std::string const& name, std::array<T, N>& v, int start = 0, int n = N) const
    {
        std::vector<T> tmp(v.begin(), v.end());
        int r = queryarr(name, tmp, start, n);
        for (size_t i = 0; i < N && i < tmp.size(); ++i) v[i] = tmp[i];
        return r;
    }

private:
    std::string full(std::string const& name) const
    {
        return m_prefix.empty() ? name : m_prefix + "." + name;
    }
    static void parse(std::string const& s, std::string& v) { v = s; }
    static void parse(std::string const& s, bool& v)
    {
        v = (s == "1" || s == "true" || s == "True" || s == "TRUE");
    }
    template <class T>
    static void parse(std::string const& s, T& v)
    {
        std::istringstream is(s);
        if constexpr (std::is_integral<T>::value)
        {
            double d;
            is >> d;
            v = T(d);
        } else
        {
            is >> v;
        }
    }
    std::string m_prefix;
};

struct PrintStream
{
    template <class T>
    PrintStream& operator<<(T const& v)
    {
        if (enabled()) std::cout << v;
        return *this;
    }
    PrintStream& operator<<(std::ostream& (*f)(std::ostream&))
    {
        if (enabled()) std::cout << f;
        return *this;
    }
    static bool& enabled()
    {
        static bool on = false;
        return on;
    }
};
inline PrintStream Print() { return PrintStream{}; }
inline PrintStream AllPrint() { return PrintStream{}; }

[[noreturn]] inline void Abort(std::string const& msg = "")
{
    throw std::runtime_error("amrex::Abort: " + msg);
}

struct Arena
{
    void* alloc(size_t n)
    {
#if defined(__HIPCC__)
        void* p = nullptr;
        if (hipMalloc(&p, n) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        return p;
#else
        return std::malloc(n);
#endif
    }
    void free(void* p)
    {
#if defined(__HIPCC__)
        (void)hipFree(p);
#else
        std::free(p);
#endif
    }
};
inline Arena* The_Arena()
{
    static Arena a;
    return &a;
}

namespace Gpu {
inline void htod_memcpy(void* dst, const void* src, size_t n)
{
#if defined(__HIPCC__)
    if (hipMemcpy(dst, src, n, hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("htod_memcpy failed");
#else
    std::memcpy(dst, src, n);
#endif
}
inline void dtoh_memcpy(void* dst, const void* src, size_t n)
{
#if defined(__HIPCC__)
    if (hipMemcpy(dst, src, n, hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("dtoh_memcpy failed");
#else
    std::memcpy(dst, src, n);
#endif
}
inline void streamSynchronize()
{
#if defined(__HIPCC__)
    (void)hipDeviceSynchronize();
#endif
}
struct HostToDevice
{
};
struct DeviceToHost
{
};
static constexpr HostToDevice hostToDevice{};
static constexpr DeviceToHost deviceToHost{};
template <class InIt, class OutIt>
void copy(HostToDevice, InIt b, InIt e, OutIt out)
{
    size_t n = size_t(e - b);
    htod_memcpy(&*out, &*b, n * sizeof(*b));
}
template <class InIt, class OutIt>
void copy(DeviceToHost, InIt b, InIt e, OutIt out)
{
    size_t n = size_t(e - b);
    dtoh_memcpy(&*out, &*b, n * sizeof(*b));
}
inline bool notInLaunchRegion() { return true; }
} // namespace Gpu

} // namespace amrex

using namespace amrex::literals;

#endif
