// prof.h -- optional per-kernel timing shared by the 1D and 3D sources: a hipEvent pair is
// recorded around each timed launch and attributed to a kernel kind; qex_profile_read sums them.
#pragma once
#include <hip/hip_runtime.h>

namespace qexprof {

// kernel kinds (index into qex_profile_read's arrays)
enum Kind {
  K_MLP = 0, K_POTENTIAL = 1, K_EIGEN = 2, K_DENSITY = 3, K_FINALIZE = 4,
  K_INT1E = 5, K_ERI = 6, K_GRID = 7, K_AO = 8, K_RHO = 9, K_VXC = 10, K_COULOMB = 11, K_FOCK = 12,
  K_EIGH = 13, K_ENERGY = 14, K_OTHER = 15, K_NN = 16, K_ADJ = 17, K_ADJ_MLP = 18, K_ADJ_EIG = 19, NKINDS = 24
};

struct Rec {
  int kind;
  hipEvent_t a, b;
};

// defined once (qex1d.hip); the library's calls are host-serialised per the threading contract
extern bool g_on;
extern unsigned g_mask;  // kinds that are timed (bit per kind)
extern Rec* g_rec;
extern int g_n, g_cap;
extern hipEvent_t* g_pool;
extern int g_pool_n, g_pool_used;

inline hipEvent_t pool_event() { return g_pool_used < g_pool_n ? g_pool[g_pool_used++] : nullptr; }

struct Scope {
  int kind;
  hipStream_t st;
  hipEvent_t b = nullptr, e = nullptr;
  Scope(int k, hipStream_t s) : kind(k), st(s) {
    if (!g_on || g_n >= g_cap || !((g_mask >> k) & 1u)) return;
    b = pool_event();
    e = pool_event();
    if (b && e) (void)hipEventRecord(b, st);
  }
  ~Scope() {
    if (b && e) {
      (void)hipEventRecord(e, st);
      g_rec[g_n++] = Rec{kind, b, e};
    }
  }
};

}  // namespace qexprof
