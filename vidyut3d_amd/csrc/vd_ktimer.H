// Per-kernel-class timing with HIP events on the launching stream.
//
// When enabled, every launch of an instrumented class is bracketed by a
// pair of events recorded on the same stream; collect() synchronises and
// sums the elapsed times.  Disabled (the default) it records nothing.
#ifndef VD_KTIMER_H_
#define VD_KTIMER_H_
#include <hip/hip_runtime.h>
#include <vector>

namespace vd {

enum KClass : int
{
    K_FLUX = 0,        // face drift flux (all directions)
    K_CELL,            // divergence + reaction source + implicit rhs
    K_SIDE,            // boundary hooks and rhs folding
    K_PRHS,            // Poisson charge-density rhs
    K_EFIELD,          // E field, E/N and potential write-back
    K_PGSRB0,          // Poisson red-black sweep, finest level
    K_PRES0,           // Poisson residual (+norm), finest level
    K_PRESTR0,         // Poisson residual+restriction, finest level
    K_PPROL0,          // Poisson prolongation onto the finest level
    K_PCOARSE,         // Poisson work on coarse levels (incl. bottom)
    K_SGSRB0,          // species red-black sweep, finest level
    K_SRES0,           // species residual (+norm), finest level
    K_SRESTR0,         // species residual+restriction, finest level
    K_SPROL0,          // species prolongation onto the finest level
    K_SCOARSE,         // species coarse levels and coefficient coarsening
    K_TSCALE,          // time-scale reduction
    K_OTHER,           // copies, scaling, LFA, fills
    K_PPRE0,           // Poisson fused pre-smoothing + residual + restriction, finest level
    K_PPOST0,          // Poisson fused prolongation + post-smoothing, finest level
    K_SPRE0,           // species fused pre-smoothing + residual + restriction, finest level
    K_SPOST0,          // species fused prolongation + post-smoothing, finest level
    K_NCLASS
};

inline const char* kclass_name(int c)
{
    static const char* n[K_NCLASS] = {"flux", "cell_source", "side_bc", "poisson_rhs", "efield",
                                      "pois_gsrb_l0", "pois_resid_l0", "pois_resid_restrict_l0", "pois_prolong_l0",
                                      "pois_coarse", "spec_gsrb_l0", "spec_resid_l0", "spec_resid_restrict_l0",
                                      "spec_prolong_l0", "spec_coarse", "time_scales", "other",
                                      "pois_pre_l0", "pois_post_l0", "spec_pre_l0", "spec_post_l0"};
    return (c >= 0 && c < K_NCLASS) ? n[c] : "?";
}

struct KTimer
{
    bool on = false;
    unsigned long long mask = ~0ULL;   // classes timed while on
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> pool;
    std::vector<int> cls;   // class of each recorded pair
    size_t used = 0;        // events in use (pairs * 2)
    double ms[K_NCLASS] = {0};
    long long count[K_NCLASS] = {0};

    hipEvent_t next()
    {
        if (used == pool.size())
        {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int c)
    {
        if (!on || !((mask >> c) & 1ULL)) return;
        cls.push_back(c);
        (void)hipEventRecord(next(), stream);
    }
    void end(int c)
    {
        if (!on || !((mask >> c) & 1ULL)) return;
        (void)hipEventRecord(next(), stream);
    }
    void collect()
    {
        if (used == 0) return;
        (void)hipStreamSynchronize(stream);
        for (size_t p = 0; p < cls.size(); ++p)
        {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, pool[2 * p], pool[2 * p + 1]);
            ms[cls[p]] += t;
            count[cls[p]] += 1;
        }
        cls.clear();
        used = 0;
    }
    void reset()
    {
        collect();
        for (int c = 0; c < K_NCLASS; ++c)
        {
            ms[c] = 0;
            count[c] = 0;
        }
    }
    ~KTimer()
    {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

struct KScope
{
    KTimer* t;
    int c;
    KScope(KTimer* t_, int c_) : t(t_), c(c_)
    {
        if (t) t->begin(c);
    }
    ~KScope()
    {
        if (t) t->end(c);
    }
};

} // namespace vd
#endif
