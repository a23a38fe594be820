// opcount.cpp -- MEASUREMENT helper (not shipped): counts the arithmetic of one RHS
// evaluation of the oracle models by compiling their C source with an operation-counting
// scalar type.  +,-,*,/ count one flop each; exp/log/pow/sqrt/floor are tallied as
// transcendentals (the algorithmic-work convention used for the cell kernels' roofline).
// Build/run: make -C oracle opcount && oracle/_ref/opcount
#include <cmath>
#include <cstdio>
#include <cstring>

struct Tally { long flops = 0, exps = 0, logs = 0, pows = 0, other = 0; };
static Tally T;

struct CD {
  double v;
  CD() : v(0) {}
  constexpr CD(double x) : v(x) {}
  CD &operator+=(const CD &o) { T.flops++; v += o.v; return *this; }
  CD &operator-=(const CD &o) { T.flops++; v -= o.v; return *this; }
  CD &operator*=(const CD &o) { T.flops++; v *= o.v; return *this; }
  CD &operator/=(const CD &o) { T.flops++; v /= o.v; return *this; }
};
static CD operator+(CD a, CD b) { T.flops++; return CD(a.v + b.v); }
static CD operator-(CD a, CD b) { T.flops++; return CD(a.v - b.v); }
static CD operator*(CD a, CD b) { T.flops++; return CD(a.v * b.v); }
static CD operator/(CD a, CD b) { T.flops++; return CD(a.v / b.v); }
static CD operator-(CD a) { return CD(-a.v); }
static bool operator<(CD a, CD b) { return a.v < b.v; }
static bool operator>(CD a, CD b) { return a.v > b.v; }
static bool operator<=(CD a, CD b) { return a.v <= b.v; }
static bool operator>=(CD a, CD b) { return a.v >= b.v; }
static bool operator==(CD a, CD b) { return a.v == b.v; }
static CD exp(CD a) { T.exps++; return CD(std::exp(a.v)); }
static CD log(CD a) { T.logs++; return CD(std::log(a.v)); }
static CD pow(CD a, CD b) { T.pows++; return CD(std::pow(a.v, b.v)); }
static CD floor(CD a) { T.other++; return CD(std::floor(a.v)); }
static CD sqrt(CD a) { T.other++; return CD(std::sqrt(a.v)); }
static CD fabs(CD a) { return CD(std::fabs(a.v)); }
static bool operator!=(CD a, CD b) { return a.v != b.v; }

// the C prototypes of oracle.h are skipped: the model sources are re-typed here
#define CEP_ORACLE_H
#define ORD_NS 52
#define ORD_NC 206
#define ORD_NA 200
#define TT_NS 19
#define TT_NC 53
#define TT_NA 70
#define ORD11_NS 41
#define ORD11_NC 144
#define ORD11_NA 198
#define ORD11_NBM 16
#include <pthread.h>
#include <time.h>
typedef struct ord11_cell_cfg_t ord11_cell_cfg_t;
#define double CD
#define memset(p, v, n) for (size_t k_ = 0; k_ < (n) / sizeof(CD); ++k_) (p)[k_] = CD(0)
#define memcpy(d, s, n) for (size_t k_ = 0; k_ < (n) / sizeof(CD); ++k_) (d)[k_] = (s)[k_]
extern "C++" {
#include "ord_cipa.c"
#include "tt06.c"
#define ORD11_ONLY_MODEL
#include "ord2011.c"
}
#undef double

int main()
{
  CD C[ORD_NC], S[ORD_NS], R[ORD_NS], A[ORD_NA];
  ord_init_consts(C, S, CD(0.0));
  T = Tally();
  ord_rates(CD(10.0), C, R, S, A);
  const Tally o = T;
  CD tc[TT_NC], ts[TT_NS], tr[TT_NS], ta[TT_NA];
  tt_init_consts(tc, ts, 0);
  T = Tally();
  tt_rates(CD(10.0), tc, tr, ts, ta, 0);
  const Tally t = T;
  T = Tally();
  tt_rush_larsen(CD(0.02), tc, tr, ts, ta);
  const Tally rl = T;
  CD oc[ORD11_NC], os[ORD11_NS], orr[ORD11_NS], oa[ORD11_NA];
  ord11_init_consts(oc, os, CD(0.0));
  T = Tally();
  ord11_rates(CD(100.0), oc, orr, os, oa);
  const Tally o11 = T;
  T = Tally();
  ord11_rush_larsen(CD(0.1), oc, orr, os, oa);
  const Tally o11rl = T;
  printf("ord11_rhs flops %ld exp %ld log %ld pow %ld other %ld\n", o11.flops, o11.exps, o11.logs, o11.pows, o11.other);
  printf("ord11_rl_update flops %ld exp %ld log %ld pow %ld other %ld\n", o11rl.flops, o11rl.exps, o11rl.logs, o11rl.pows, o11rl.other);
  printf("ord_rhs flops %ld exp %ld log %ld pow %ld other %ld\n", o.flops, o.exps, o.logs, o.pows, o.other);
  printf("tt_rhs flops %ld exp %ld log %ld pow %ld other %ld\n", t.flops, t.exps, t.logs, t.pows, t.other);
  printf("tt_rl_update flops %ld exp %ld log %ld pow %ld other %ld\n", rl.flops, rl.exps, rl.logs, rl.pows, rl.other);
  return 0;
}
