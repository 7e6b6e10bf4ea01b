// cpu_ref — CPU restatement of Grid's Wilson / Mobius-DWF Dirac operator and
// even-odd CG, used ONLY as test infrastructure: the parity checker for the HIP
// library (tests/), the check inside __graft_entry__.smoke(), and the CPU
// baseline leg of bench.py.  No product code path links or calls this file.
//
// Semantics follow paboyle/Grid (paths relative to the Grid source tree):
//   spin projection / reconstruction      Grid/qcd/spin/TwoSpinor.h:67-323
//   Dhop site kernel (leg order, projectors, link choice)
//                                         Grid/qcd/action/fermion/implementation/WilsonKernelsImplementation.h:112-163
//   multLink                              Grid/qcd/action/fermion/WilsonImpl.h:83-101
//   DoubleStore (-1/2 scaling, U^dag(x-mu), boundary phases)
//                                         Grid/qcd/action/fermion/WilsonImpl.h:127-171,
//                                         Grid/qcd/action/fermion/implementation/WilsonFermionImplementation.h:82-107
//   Wilson M/Meooe/Mooee/MooeeInv         WilsonFermionImplementation.h:113-176
//   5D hop with shared links, DW          WilsonFermion5DImplementation.h:414-452
//   Cayley M5D / M5Ddag / MooeeInv(Dag)   Grid/qcd/action/fermion/CayleyFermion5Dcache.h:41-230
//   Cayley compositions                   CayleyFermion5DImplementation.h:155-329
//   Cayley coefficients + LDU             CayleyFermion5DImplementation.h:410-535 (Mobius tanh: omega_s = 1,
//                                         Grid/algorithms/approx/Zolotarev.cc:473)
//   Schur operator                        Grid/algorithms/LinearOperator.h:286-349
//   Red-black source / solution           Grid/algorithms/iterative/SchurRedBlack.h:236-291,385-425
//   CG loop                               Grid/algorithms/iterative/ConjugateGradient.h:68-257
//   Reductions (per-thread partial sums, serial combine in double)
//                                         Grid/lattice/Lattice_reduction.h:71-100,255-287
//   Checkerboard indexing                 Grid/cartesian/Cartesian_red_black.h:39-95, Grid/qcd/utils/SpaceTimeGrid.cc:66-78
//
// Host layouts (Grid lexicographic order, complex as interleaved re,im doubles):
//   gauge      U  [site4][mu][a][b]                  site4 = x + Lx*(y + Ly*(z + Lz*t))
//   doubled    Uds[site4][8][a][b]                   point p: mu=p%4, +mu for p<4, -mu for p>=4
//   fermion    psi[site][s][spin][colour]            full: site4; half: h = (x/2) + (Lx/2)*(y + Ly*(z + Lz*t))
//   5D fields keep s innermost, as Grid's FiveDim grids do.
#include <complex>
#include <vector>
#include <cmath>
#include <cstring>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <omp.h>

typedef std::complex<double> cd;

namespace {

struct Geom {
  int L[4];
  int Ls;
  int64_t V4, H4;
  explicit Geom(const int *d, int ls) {
    for (int i = 0; i < 4; i++) L[i] = d[i];
    Ls = ls;
    V4 = (int64_t)L[0] * L[1] * L[2] * L[3];
    H4 = V4 / 2;
  }
  int64_t lex(const int x[4]) const { return x[0] + (int64_t)L[0] * (x[1] + (int64_t)L[1] * (x[2] + (int64_t)L[2] * x[3])); }
  void coor(int64_t idx, int x[4]) const {
    for (int d = 0; d < 4; d++) { x[d] = (int)(idx % L[d]); idx /= L[d]; }
  }
  static int parity(const int x[4]) { return (x[0] + x[1] + x[2] + x[3]) & 1; }
  int64_t half(const int x[4]) const { return (x[0] >> 1) + (int64_t)(L[0] / 2) * (x[1] + (int64_t)L[1] * (x[2] + (int64_t)L[2] * x[3])); }
  void coor_half(int64_t h, int cb, int x[4]) const {
    int xh = (int)(h % (L[0] / 2)); h /= (L[0] / 2);
    x[1] = (int)(h % L[1]); h /= L[1];
    x[2] = (int)(h % L[2]); h /= L[2];
    x[3] = (int)h;
    x[0] = 2 * xh + ((cb + x[1] + x[2] + x[3]) & 1);
  }
  void neighbour(const int x[4], int mu, int disp, int y[4]) const {
    for (int d = 0; d < 4; d++) y[d] = x[d];
    y[mu] = (x[mu] + disp + L[mu]) % L[mu];
  }
};

struct Spinor { cd f[4][3]; };
struct Half { cd h[2][3]; };

const cd I(0.0, 1.0);

// (1 + gamma_mu) and (1 - gamma_mu) projections, TwoSpinor.h:75-133
inline void proj(Half &h, const cd *f /*[4][3]*/, int mu, int plus) {
  const cd *f0 = f, *f1 = f + 3, *f2 = f + 6, *f3 = f + 9;
  for (int c = 0; c < 3; c++) {
    switch (mu * 2 + (plus ? 0 : 1)) {
      case 0: h.h[0][c] = f0[c] + I * f3[c]; h.h[1][c] = f1[c] + I * f2[c]; break;  // Xp
      case 1: h.h[0][c] = f0[c] - I * f3[c]; h.h[1][c] = f1[c] - I * f2[c]; break;  // Xm
      case 2: h.h[0][c] = f0[c] - f3[c];     h.h[1][c] = f1[c] + f2[c];     break;  // Yp
      case 3: h.h[0][c] = f0[c] + f3[c];     h.h[1][c] = f1[c] - f2[c];     break;  // Ym
      case 4: h.h[0][c] = f0[c] + I * f2[c]; h.h[1][c] = f1[c] - I * f3[c]; break;  // Zp
      case 5: h.h[0][c] = f0[c] - I * f2[c]; h.h[1][c] = f1[c] + I * f3[c]; break;  // Zm
      case 6: h.h[0][c] = f0[c] + f2[c];     h.h[1][c] = f1[c] + f3[c];     break;  // Tp
      case 7: h.h[0][c] = f0[c] - f2[c];     h.h[1][c] = f1[c] - f3[c];     break;  // Tm
    }
  }
}

// accumRecon{X,Y,Z,T}{p,m}, TwoSpinor.h:193-323
inline void accum_recon(Spinor &r, const Half &h, int mu, int plus) {
  for (int c = 0; c < 3; c++) {
    const cd h0 = h.h[0][c], h1 = h.h[1][c];
    r.f[0][c] += h0;
    r.f[1][c] += h1;
    switch (mu * 2 + (plus ? 0 : 1)) {
      case 0: r.f[2][c] -= I * h1; r.f[3][c] -= I * h0; break;
      case 1: r.f[2][c] += I * h1; r.f[3][c] += I * h0; break;
      case 2: r.f[2][c] += h1;     r.f[3][c] -= h0;     break;
      case 3: r.f[2][c] -= h1;     r.f[3][c] += h0;     break;
      case 4: r.f[2][c] -= I * h0; r.f[3][c] += I * h1; break;
      case 5: r.f[2][c] += I * h0; r.f[3][c] -= I * h1; break;
      case 6: r.f[2][c] += h0;     r.f[3][c] += h1;     break;
      case 7: r.f[2][c] -= h0;     r.f[3][c] -= h1;     break;
    }
  }
}

// U * chi for both spin rows, WilsonImpl.h:83-101
inline void mult_link(Half &out, const cd *U /*[3][3]*/, const Half &in) {
  for (int s = 0; s < 2; s++)
    for (int a = 0; a < 3; a++)
      out.h[s][a] = U[a * 3 + 0] * in.h[s][0] + U[a * 3 + 1] * in.h[s][1] + U[a * 3 + 2] * in.h[s][2];
}

// 5D hop. Output sites: all (cb_out<0) or one checkerboard (cb_out = 0 even, 1 odd).
// WilsonKernelsImplementation.h:112-163 (leg order Xm,Ym,Zm,Tm,Xp,Yp,Zp,Tp for DaggerNo;
// Xp..Tp, Xm..Tm for DaggerYes); 5D: WilsonFermion5DImplementation.h:414-452, sF = s + Ls*sU.
void dhop(const Geom &g, const cd *Uds, const cd *in, cd *out, int cb_out, int dag) {
  const int Ls = g.Ls;
  const int64_t nout = (cb_out < 0) ? g.V4 : g.H4;
#pragma omp parallel for schedule(static)
  for (int64_t so = 0; so < nout; so++) {
    int x[4];
    if (cb_out < 0) g.coor(so, x); else g.coor_half(so, cb_out, x);
    const int64_t sU = g.lex(x);
    // neighbour site indices for the 8 points (in the input field's indexing)
    int64_t nb[8];
    for (int p = 0; p < 8; p++) {
      int y[4];
      g.neighbour(x, p % 4, p < 4 ? +1 : -1, y);
      nb[p] = (cb_out < 0) ? g.lex(y) : g.half(y);
    }
    for (int s = 0; s < Ls; s++) {
      Spinor r;
      for (int a = 0; a < 4; a++) for (int c = 0; c < 3; c++) r.f[a][c] = 0.0;
      for (int leg = 0; leg < 8; leg++) {
        // DaggerNo: first the backward points (Xm..Tm) with (1+gamma), then forward (Xp..Tp) with (1-gamma)
        // DaggerYes: forward points with (1+gamma), then backward with (1-gamma)
        int p, plus;
        if (!dag) { p = (leg < 4) ? leg + 4 : leg - 4; plus = (leg < 4); }
        else      { p = (leg < 4) ? leg : leg; plus = (leg < 4); }
        const int mu = p % 4;
        const cd *psi = in + (nb[p] * Ls + s) * 12;
        Half chi, Uchi;
        proj(chi, psi, mu, plus);
        mult_link(Uchi, Uds + (sU * 8 + p) * 9, chi);
        accum_recon(r, Uchi, mu, plus);
      }
      const int64_t sF = (cb_out < 0 ? so : so) * Ls + s;
      cd *o = out + sF * 12;
      for (int a = 0; a < 4; a++) for (int c = 0; c < 3; c++) o[a * 3 + c] = r.f[a][c];
    }
  }
}

// Cayley M5D kernels, CayleyFermion5Dcache.h:41-115.  P+ = spins 0,1; P- = spins 2,3 (spProj5p/m).
void m5d(const Geom &g, int64_t n4, const cd *psi, const cd *phi, cd *chi,
         const double *lower, const double *diag, const double *upper, int dag) {
  const int Ls = g.Ls;
  std::vector<cd> tmp;
  const cd *phiv = phi;
  if (phi == chi) { tmp.assign(phi, phi + n4 * Ls * 12); phiv = tmp.data(); }
  const cd *psiv = psi;
  std::vector<cd> tmp2;
  if (psi == chi) { tmp2.assign(psi, psi + n4 * Ls * 12); psiv = tmp2.data(); }
#pragma omp parallel for schedule(static)
  for (int64_t sss = 0; sss < n4 * Ls; sss++) {
    const int64_t s = sss % Ls, ss = sss - s;
    const cd *pu = psiv + (ss + (s + 1) % Ls) * 12;
    const cd *pl = psiv + (ss + (s + Ls - 1) % Ls) * 12;
    const cd *ph = phiv + sss * 12;
    cd *o = chi + sss * 12;
    for (int sp = 0; sp < 4; sp++)
      for (int c = 0; c < 3; c++) {
        const int k = sp * 3 + c;
        const bool upper_half = sp < 2;  // P+ keeps spins 0,1
        cd up, lo;
        if (!dag) { up = upper_half ? cd(0) : pu[k]; lo = upper_half ? pl[k] : cd(0); }   // P- psi(s+1), P+ psi(s-1)
        else      { up = upper_half ? pu[k] : cd(0); lo = upper_half ? cd(0) : pl[k]; }   // P+ psi(s+1), P- psi(s-1)
        o[k] = diag[s] * ph[k] + upper[s] * up + lower[s] * lo;
      }
  }
}

struct Cayley {
  int Ls;
  double mass, M5;
  std::vector<double> bs, cs, bee, cee, beo, ceo, dee, lee, leem, uee, ueem;
  // CayleyFermion5DImplementation.h:410-535 with omega_s = 1 (Mobius tanh)
  Cayley(int ls, double m, double m5, double b, double c) : Ls(ls), mass(m), M5(m5) {
    bs.resize(Ls); cs.resize(Ls); bee.resize(Ls); cee.resize(Ls); beo.resize(Ls); ceo.resize(Ls);
    std::vector<double> aee(Ls);
    const double bpc = b + c, bmc = b - c;
    for (int i = 0; i < Ls; i++) {
      const double omega = 1.0;
      bs[i] = 0.5 * (bpc / omega + bmc);
      cs[i] = 0.5 * (bpc / omega - bmc);
    }
    for (int i = 0; i < Ls; i++) {
      bee[i] = bs[i] * (4.0 - M5) + 1.0;
      cee[i] = 1.0 - cs[i] * (4.0 - M5);
      beo[i] = bs[i];
      ceo[i] = -cs[i];
      aee[i] = cee[i];
    }
    dee.resize(Ls); lee.resize(Ls); leem.resize(Ls); uee.resize(Ls); ueem.resize(Ls);
    const double mass_plus = mass, mass_minus = mass;
    for (int i = 0; i < Ls; i++) {
      dee[i] = bee[i];
      if (i < Ls - 1) {
        lee[i] = -cee[i + 1] / bee[i];
        leem[i] = mass_minus * cee[Ls - 1] / bee[0];
        for (int j = 0; j < i; j++) leem[i] *= aee[j] / bee[j + 1];
        uee[i] = -aee[i] / bee[i];
        ueem[i] = mass_plus;
        for (int j = 1; j <= i; j++) ueem[i] *= cee[j] / bee[j];
        ueem[i] *= aee[0] / bee[0];
      } else {
        lee[i] = 0; leem[i] = 0; uee[i] = 0; ueem[i] = 0;
      }
    }
    double delta_d = mass_minus * cee[Ls - 1];
    for (int j = 0; j < Ls - 1; j++) delta_d *= cee[j] / bee[j];
    dee[Ls - 1] += delta_d;
  }
};

// MooeeInv / MooeeInvDag LDU sweeps, CayleyFermion5Dcache.h:117-230
void mooee_inv(const Geom &g, const Cayley &C, int64_t n4, const cd *psi, cd *chi, int dag) {
  const int Ls = g.Ls;
#pragma omp parallel for schedule(static)
  for (int64_t s4 = 0; s4 < n4; s4++) {
    const int64_t ss = s4 * Ls;
    for (int k = 0; k < 12; k++) {
      const bool up = (k < 6);  // spins 0,1
      // With P+ taking the upper spins, each (spin,colour) component runs its own scalar recurrence.
      // Non-dag: tmp = P+ res (upper comps), acc accumulates P- res (lower comps).
      // Dag: roles of P+ and P- exchanged, coefficients lee<->uee, leem<->ueem (real here, conj no-op).
      const bool carry = dag ? !up : up;  // component carried by tmp through the forward sweep
      cd tmp = 0.0, acc = 0.0, res;
      const double *l1 = dag ? C.uee.data() : C.lee.data();
      const double *l2 = dag ? C.ueem.data() : C.leem.data();
      const double *u1 = dag ? C.lee.data() : C.uee.data();
      const double *u2 = dag ? C.leem.data() : C.ueem.data();
      // forward sweep
      res = psi[ss * 12 + k];
      if (carry) tmp = res; else acc = l2[0] * res;
      chi[ss * 12 + k] = res;
      for (int s = 1; s < Ls - 1; s++) {
        res = psi[(ss + s) * 12 + k];
        if (carry) res -= l1[s - 1] * tmp;
        if (carry) tmp = res; else acc += l2[s] * res;
        chi[(ss + s) * 12 + k] = res;
      }
      res = psi[(ss + Ls - 1) * 12 + k];
      if (carry) res -= l1[Ls - 2] * tmp;
      if (!carry) res -= acc;
      // backward sweep
      res = (1.0 / C.dee[Ls - 1]) * res;
      chi[(ss + Ls - 1) * 12 + k] = res;
      // After the last site: acc = P_carry-complement? Grid: non-dag acc = P+ res(Ls-1), tmp = P- res(Ls-1)
      cd accb = carry ? res : cd(0.0);
      cd tmpb = carry ? cd(0.0) : res;
      for (int s = Ls - 2; s >= 0; s--) {
        res = (1.0 / C.dee[s]) * chi[(ss + s) * 12 + k] - u1[s] * tmpb - u2[s] * accb;
        tmpb = carry ? cd(0.0) : res;
        chi[(ss + s) * 12 + k] = res;
      }
    }
  }
}

void axpby(int64_t n, double a, const cd *x, double b, const cd *y, cd *z) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) z[i] = a * x[i] + b * y[i];
}

// Lattice_reduction.h:71-100: per-thread partial sums over a static split, serial combine.
cd inner(int64_t n, const cd *x, const cd *y) {
  const int nt = omp_get_max_threads();
  std::vector<cd> part(nt, 0.0);
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num();
    const int64_t chunk = (n + nt - 1) / nt;
    const int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    cd s = 0.0;
    for (int64_t i = b; i < e; i++) s += std::conj(x[i]) * y[i];
    part[t] = s;
  }
  cd s = 0.0;
  for (int t = 0; t < nt; t++) s += part[t];
  return s;
}
double norm2(int64_t n, const cd *x) { return std::real(inner(n, x, x)); }

enum OpKind { OP_M = 0, OP_MDAG, OP_MEOOE, OP_MEOOEDAG, OP_MOOEE, OP_MOOEEDAG, OP_MOOEEINV, OP_MOOEEINVDAG,
              OP_M5D, OP_M5DDAG, OP_MEOOE5D, OP_MEOOEDAG5D, OP_DHOP, OP_DHOPDAG };

struct Fermion {
  // kind 0 = Wilson (Ls copies of the 4D operator: multi-RHS when Ls>1), 1 = Mobius/Cayley
  int kind;
  Geom g;
  const cd *Uds;
  double mass;
  Cayley *C;
  Fermion(int k, const Geom &geom, const cd *U, double m, Cayley *c) : kind(k), g(geom), Uds(U), mass(m), C(c) {}
  int64_t nhalf() const { return g.H4 * g.Ls * 12; }
  int64_t nfull() const { return g.V4 * g.Ls * 12; }

  // M5D family, CayleyFermion5DImplementation.h:155-271
  void cay_m5d(int64_t n4, const cd *psi, cd *chi, int op) const {
    const int Ls = g.Ls;
    std::vector<double> lower(Ls), diag(Ls), upper(Ls);
    const Cayley &c = *C;
    int dag = 0;
    const cd *phi = psi;
    switch (op) {
      case OP_M5D:  // diag 1, upper -1 (corner mass), lower -1 (corner mass); phi = chi
        for (int s = 0; s < Ls; s++) { diag[s] = 1; upper[s] = -1; lower[s] = -1; }
        upper[Ls - 1] = mass; lower[0] = mass; phi = chi; break;
      case OP_M5DDAG:
        for (int s = 0; s < Ls; s++) { diag[s] = 1; upper[s] = -1; lower[s] = -1; }
        upper[Ls - 1] = mass; lower[0] = mass; phi = chi; dag = 1; break;
      case OP_MEOOE5D:
        for (int s = 0; s < Ls; s++) { diag[s] = c.bs[s]; upper[s] = c.cs[s]; lower[s] = c.cs[s]; }
        upper[Ls - 1] = -mass * upper[Ls - 1]; lower[0] = -mass * lower[0]; break;
      case OP_MEOOEDAG5D:
        for (int s = 0; s < Ls; s++) {
          diag[s] = c.bs[s];
          if (s == 0) { upper[s] = c.cs[s + 1]; lower[s] = -mass * c.cs[Ls - 1]; }
          else if (s == Ls - 1) { upper[s] = -mass * c.cs[0]; lower[s] = c.cs[s - 1]; }
          else { upper[s] = c.cs[s + 1]; lower[s] = c.cs[s - 1]; }
        }
        dag = 1; break;
      case OP_MOOEE:
        for (int s = 0; s < Ls; s++) { diag[s] = c.bee[s]; upper[s] = -c.cee[s]; lower[s] = -c.cee[s]; }
        upper[Ls - 1] = -mass * upper[Ls - 1]; lower[0] = -mass * lower[0]; break;
      case OP_MOOEEDAG:
        for (int s = 0; s < Ls; s++) {
          diag[s] = c.bee[s];
          if (s == 0) { upper[s] = -c.cee[s + 1]; lower[s] = mass * c.cee[Ls - 1]; }
          else if (s == Ls - 1) { upper[s] = mass * c.cee[0]; lower[s] = -c.cee[s - 1]; }
          else { upper[s] = -c.cee[s + 1]; lower[s] = -c.cee[s - 1]; }
        }
        dag = 1; break;
      default: abort();
    }
    m5d(g, n4, psi, phi, chi, lower.data(), diag.data(), upper.data(), dag);
  }

  // out(cb_out) = hop(in(!cb_out)); cb_out<0 for full lattice
  void hop(const cd *in, cd *out, int cb_out, int dag) const { dhop(g, Uds, in, out, cb_out, dag); }

  // Full-lattice M / Mdag
  void M(const cd *in, cd *out, int dag) const {
    const int64_t n = nfull();
    if (kind == 0) {  // WilsonFermionImplementation.h:113-127: Dhop + (4+m)
      hop(in, out, -1, dag);
      axpby(n, 1.0, out, 4.0 + mass, in, out);
      return;
    }
    std::vector<cd> Din(n);
    const double diagDW = 4.0 - C->M5;  // WilsonFermion5DImplementation.h:446-452
    if (!dag) {  // CayleyFermion5DImplementation.h:273-285
      cay_m5d(g.V4, in, Din.data(), OP_MEOOE5D);
      hop(Din.data(), out, -1, 0);
      axpby(n, 1.0, out, diagDW, Din.data(), out);  // DW
      axpby(n, 1.0, out, 1.0, in, out);
      cay_m5d(g.V4, in, out, OP_M5D);
    } else {  // :287-302
      hop(in, Din.data(), -1, 1);
      axpby(n, 1.0, Din.data(), diagDW, in, Din.data());
      cay_m5d(g.V4, Din.data(), out, OP_MEOOEDAG5D);
      cay_m5d(g.V4, in, out, OP_M5DDAG);
      axpby(n, 1.0, out, 1.0, in, out);
    }
  }
  // Meooe: psi on cb_in, output on !cb_in
  void Meooe(const cd *in, cd *out, int cb_in, int dag) const {
    const int64_t n = nhalf();
    if (kind == 0) { hop(in, out, 1 - cb_in, dag); return; }
    std::vector<cd> tmp(n);
    if (!dag) {  // CayleyFermion5DImplementation.h:305-317
      cay_m5d(g.H4, in, tmp.data(), OP_MEOOE5D);
      hop(tmp.data(), out, 1 - cb_in, 0);
    } else {  // :319-329
      hop(in, tmp.data(), 1 - cb_in, 1);
      cay_m5d(g.H4, tmp.data(), out, OP_MEOOEDAG5D);
    }
  }
  void Mooee(const cd *in, cd *out, int dag) const {
    const int64_t n = nhalf();
    if (kind == 0) { axpby(n, 4.0 + mass, in, 0.0, in, out); return; }
    cay_m5d(g.H4, in, out, dag ? OP_MOOEEDAG : OP_MOOEE);
  }
  void MooeeInv(const cd *in, cd *out, int dag) const {
    const int64_t n = nhalf();
    if (kind == 0) { axpby(n, 1.0 / (4.0 + mass), in, 0.0, in, out); return; }
    mooee_inv(g, *C, g.H4, in, out, dag);
  }
  // SchurDiagMooeeOperator, LinearOperator.h:325-349
  void Mpc(const cd *in, cd *out, int cb, int dag) const {
    const int64_t n = nhalf();
    std::vector<cd> tmp(n), o2(n);
    if (!dag) {
      Meooe(in, tmp.data(), cb, 0);
      MooeeInv(tmp.data(), o2.data(), 0);
      Meooe(o2.data(), tmp.data(), 1 - cb, 0);
      Mooee(in, out, 0);
    } else {
      Meooe(in, tmp.data(), cb, 1);
      MooeeInv(tmp.data(), o2.data(), 1);
      Meooe(o2.data(), tmp.data(), 1 - cb, 1);
      Mooee(in, out, 1);
    }
    axpby(n, 1.0, out, -1.0, tmp.data(), out);
  }
  void MpcDagMpc(const cd *in, cd *out, int cb) const {
    std::vector<cd> tmp(nhalf());
    Mpc(in, tmp.data(), cb, 0);
    Mpc(tmp.data(), out, cb, 1);
  }
};

// ConjugateGradient::operator(), ConjugateGradient.h:68-257, on HermOp = MpcDagMpc on checkerboard cb.
int cg(const Fermion &D, int cb, const cd *src, cd *psi, double tol, int maxit, int *iters, double *true_res) {
  const int64_t n = D.nhalf();
  std::vector<cd> p(n), mmp(n), r(n);
  double cp, c, a, d, b, ssq;
  ssq = norm2(n, src);
  double guess = norm2(n, psi);
  if (std::isnan(guess)) return -1;
  if (guess == 0.0) {
    std::copy(src, src + n, r.begin());
    p = r;
    a = ssq;
  } else {
    D.MpcDagMpc(psi, mmp.data(), cb);
    axpby(n, 1.0, src, -1.0, mmp.data(), r.data());
    p = r;
    a = norm2(n, p.data());
  }
  cp = a;
  if (ssq == 0.) {
    std::fill(psi, psi + n, cd(0));
    *iters = 1; *true_res = 0.; return 0;
  }
  const double rsq = tol * tol * ssq;
  if (cp <= rsq) { *true_res = std::sqrt(a / ssq); *iters = 0; return 0; }
  for (int k = 1; k <= maxit; k++) {
    c = cp;
    D.MpcDagMpc(p.data(), mmp.data(), cb);
    d = std::real(inner(n, p.data(), mmp.data()));
    a = c / d;
    axpby(n, -a, mmp.data(), 1.0, r.data(), r.data());
    cp = norm2(n, r.data());
    b = cp / c;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      psi[i] = a * p[i] + psi[i];
      p[i] = b * p[i] + r[i];
    }
    if (cp <= rsq) {
      D.MpcDagMpc(psi, mmp.data(), cb);
      axpby(n, 1.0, mmp.data(), -1.0, src, p.data());
      *true_res = std::sqrt(norm2(n, p.data())) / std::sqrt(norm2(n, src));
      *iters = k;
      return 0;
    }
  }
  *iters = maxit;
  *true_res = -1;
  return 1;
}

}  // namespace

extern "C" {

typedef struct {
  int dims[4];
  int Ls;
  int kind;  // 0 Wilson, 1 Mobius
  double mass, M5, b, c;
  double phase_re[4], phase_im[4];  // boundary phases (1,0 = periodic)
} oracle_params;

int oracle_num_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) { omp_set_num_threads(n); }

// DoubleStore with -1/2 scaling: Uds[x][mu] = -1/2 U_mu(x) (phase at x_mu = L-1),
// Uds[x][mu+4] = -1/2 U_mu(x-mu)^dag (conj phase at x_mu = 0).  WilsonImpl.h:127-171; WilsonFermionImplementation.h:102.
void oracle_double_store(const oracle_params *P, const double *U_, double *Uds_) {
  Geom g(P->dims, 1);
  const cd *U = reinterpret_cast<const cd *>(U_);
  cd *Uds = reinterpret_cast<cd *>(Uds_);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < g.V4; s++) {
    int x[4];
    g.coor(s, x);
    for (int mu = 0; mu < 4; mu++) {
      const cd ph(P->phase_re[mu], P->phase_im[mu]);
      const cd *u = U + (s * 4 + mu) * 9;
      cd *o = Uds + (s * 8 + mu) * 9;
      const cd f = (x[mu] == g.L[mu] - 1) ? ph : cd(1.0);
      for (int k = 0; k < 9; k++) o[k] = -0.5 * (f * u[k]);
      int y[4];
      g.neighbour(x, mu, -1, y);
      const cd *ub = U + (g.lex(y) * 4 + mu) * 9;
      cd *ob = Uds + (s * 8 + mu + 4) * 9;
      const cd fb = (x[mu] == 0) ? std::conj(ph) : cd(1.0);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) ob[a * 3 + b] = -0.5 * (fb * std::conj(ub[b * 3 + a]));
    }
  }
}

// out = Dhop(in); cb_out = -1 full lattice, 0 = DhopOE (out even), 1 = DhopEO... NB Grid's DhopEO writes Even
// from Odd input: use cb_out = 0 for DhopEO, cb_out = 1 for DhopOE.
void oracle_dhop(const oracle_params *P, const double *Uds, const double *in, double *out, int cb_out, int dag) {
  Geom g(P->dims, P->Ls);
  dhop(g, reinterpret_cast<const cd *>(Uds), reinterpret_cast<const cd *>(in), reinterpret_cast<cd *>(out), cb_out, dag);
}

// Generic operator application.  op: OpKind.  cb = checkerboard of the input for half-lattice ops.
int oracle_apply(const oracle_params *P, const double *Uds, const double *in_, double *out_, int op, int cb) {
  Geom g(P->dims, P->Ls);
  Cayley *C = nullptr;
  if (P->kind == 1) C = new Cayley(P->Ls, P->mass, P->M5, P->b, P->c);
  Fermion D(P->kind, g, reinterpret_cast<const cd *>(Uds), P->mass, C);
  const cd *in = reinterpret_cast<const cd *>(in_);
  cd *out = reinterpret_cast<cd *>(out_);
  int rc = 0;
  switch (op) {
    case OP_M: D.M(in, out, 0); break;
    case OP_MDAG: D.M(in, out, 1); break;
    case OP_MEOOE: D.Meooe(in, out, cb, 0); break;
    case OP_MEOOEDAG: D.Meooe(in, out, cb, 1); break;
    case OP_MOOEE: D.Mooee(in, out, 0); break;
    case OP_MOOEEDAG: D.Mooee(in, out, 1); break;
    case OP_MOOEEINV: D.MooeeInv(in, out, 0); break;
    case OP_MOOEEINVDAG: D.MooeeInv(in, out, 1); break;
    case OP_M5D: case OP_M5DDAG: case OP_MEOOE5D: case OP_MEOOEDAG5D:
      if (P->kind != 1) { rc = -1; break; }
      if (op == OP_M5D || op == OP_M5DDAG) {  // chi input holds phi; Grid's M5D(psi,chi) uses chi as phi
        std::memcpy(out_, in_, sizeof(cd) * g.H4 * g.Ls * 12);
      }
      D.cay_m5d(g.H4, in, out, op);
      break;
    case OP_DHOP: D.hop(in, out, 1 - cb, 0); break;
    case OP_DHOPDAG: D.hop(in, out, 1 - cb, 1); break;
    default: rc = -1;
  }
  delete C;
  return rc;
}

// which: 0 Mpc, 1 MpcDag, 2 MpcDagMpc; cb = checkerboard of the field (Odd = 1 for Grid's solves)
void oracle_schur(const oracle_params *P, const double *Uds, const double *in, double *out, int which, int cb) {
  Geom g(P->dims, P->Ls);
  Cayley *C = (P->kind == 1) ? new Cayley(P->Ls, P->mass, P->M5, P->b, P->c) : nullptr;
  Fermion D(P->kind, g, reinterpret_cast<const cd *>(Uds), P->mass, C);
  const cd *i = reinterpret_cast<const cd *>(in);
  cd *o = reinterpret_cast<cd *>(out);
  if (which == 0) D.Mpc(i, o, cb, 0);
  else if (which == 1) D.Mpc(i, o, cb, 1);
  else D.MpcDagMpc(i, o, cb);
  delete C;
}

// ImportPhysicalFermionSource (CayleyFermion5DImplementation.h:114-129): InsertSlice at s=0 and
// s=Ls-1, P+ on slice 0 and P- on slice Ls-1 (LinalgUtils.h:140-183), then Dminus (:130-139):
// chi_s = psi_s - cs_s DW(psi)_s with DW = Dhop + (4-M5) (WilsonFermion5DImplementation.h:447-452).
void oracle_import_physical(const oracle_params *P, const double *Uds, const double *in4d_, double *out5d_) {
  Geom g(P->dims, P->Ls);
  Cayley C(P->Ls, P->mass, P->M5, P->b, P->c);
  const int Ls = g.Ls;
  const cd *in = reinterpret_cast<const cd *>(in4d_);
  cd *out = reinterpret_cast<cd *>(out5d_);
  std::vector<cd> tmp(g.V4 * Ls * 12, 0.0), dw(g.V4 * Ls * 12);
  for (int64_t s4 = 0; s4 < g.V4; s4++)
    for (int k = 0; k < 12; k++) {
      const bool up = k < 6;
      tmp[(s4 * Ls + 0) * 12 + k] = up ? in[s4 * 12 + k] : cd(0);
      tmp[(s4 * Ls + Ls - 1) * 12 + k] = up ? cd(0) : in[s4 * 12 + k];
    }
  dhop(g, reinterpret_cast<const cd *>(Uds), tmp.data(), dw.data(), -1, 0);
  axpby((int64_t)dw.size(), 1.0, dw.data(), 4.0 - P->M5, tmp.data(), dw.data());
  for (int64_t s4 = 0; s4 < g.V4; s4++)
    for (int s = 0; s < Ls; s++)
      for (int k = 0; k < 12; k++) {
        const int64_t i = (s4 * Ls + s) * 12 + k;
        out[i] = 1.0 * tmp[i] + (-C.cs[s]) * dw[i];
      }
}

// ExportPhysicalFermionSolution (CayleyFermion5DImplementation.h:57-67): P- psi(0) + P+ psi(Ls-1).
void oracle_export_physical(const oracle_params *P, const double *sol5d_, double *out4d_) {
  Geom g(P->dims, P->Ls);
  const int Ls = g.Ls;
  const cd *sol = reinterpret_cast<const cd *>(sol5d_);
  cd *out = reinterpret_cast<cd *>(out4d_);
  for (int64_t s4 = 0; s4 < g.V4; s4++)
    for (int k = 0; k < 12; k++) {
      const bool up = k < 6;
      const cd lo = up ? cd(0) : sol[(s4 * Ls + 0) * 12 + k];
      const cd hi = up ? sol[(s4 * Ls + Ls - 1) * 12 + k] : cd(0);
      out[s4 * 12 + k] = lo + hi;
    }
}

// Cayley coefficient vectors in order bs,cs,bee,cee,dee,lee,leem,uee,ueem (each Ls doubles)
void oracle_cayley_coeffs(const oracle_params *P, double *out) {
  Cayley C(P->Ls, P->mass, P->M5, P->b, P->c);
  const std::vector<double> *v[9] = {&C.bs, &C.cs, &C.bee, &C.cee, &C.dee, &C.lee, &C.leem, &C.uee, &C.ueem};
  for (int k = 0; k < 9; k++) std::memcpy(out + k * P->Ls, v[k]->data(), sizeof(double) * P->Ls);
}

// CG on MpcDagMpc (odd checkerboard by default).  Returns 0 converged, 1 not converged.
int oracle_cg(const oracle_params *P, const double *Uds, const double *src, double *psi, int cb, double tol,
              int maxit, int *iters, double *true_res) {
  Geom g(P->dims, P->Ls);
  Cayley *C = (P->kind == 1) ? new Cayley(P->Ls, P->mass, P->M5, P->b, P->c) : nullptr;
  Fermion D(P->kind, g, reinterpret_cast<const cd *>(Uds), P->mass, C);
  const int rc = cg(D, cb, reinterpret_cast<const cd *>(src), reinterpret_cast<cd *>(psi), tol, maxit, iters, true_res);
  delete C;
  return rc;
}

// SchurRedBlackDiagMooeeSolve on a full-lattice source (SchurRedBlack.h:236-291,385-425).
// src, sol: full 5D (or 4D) fields in lexicographic site order.  Reports the CG iterations,
// the CG true residual and the unpreconditioned residual |M sol - src| / |src|.
int oracle_schur_solve(const oracle_params *P, const double *Uds, const double *src_, double *sol_, double tol,
                       int maxit, int *iters, double *cg_true_res, double *unprec_res) {
  Geom g(P->dims, P->Ls);
  Cayley *C = (P->kind == 1) ? new Cayley(P->Ls, P->mass, P->M5, P->b, P->c) : nullptr;
  Fermion D(P->kind, g, reinterpret_cast<const cd *>(Uds), P->mass, C);
  const cd *src = reinterpret_cast<const cd *>(src_);
  cd *sol = reinterpret_cast<cd *>(sol_);
  const int Ls = g.Ls;
  const int64_t nh = D.nhalf(), nf = D.nfull();
  std::vector<cd> se(nh), so(nh), tmp(nh), Mtmp(nh), solo(nh, 0.0), sole(nh);
  // pickCheckerboard
  for (int64_t s4 = 0; s4 < g.V4; s4++) {
    int x[4];
    g.coor(s4, x);
    const int cb = Geom::parity(x);
    const int64_t h = g.half(x);
    cd *dst = (cb == 0 ? se.data() : so.data()) + h * Ls * 12;
    std::memcpy(dst, src + s4 * Ls * 12, sizeof(cd) * Ls * 12);
  }
  const int Even = 0, Odd = 1;
  D.MooeeInv(se.data(), tmp.data(), 0);
  D.Meooe(tmp.data(), Mtmp.data(), Even, 0);
  axpby(nh, 1.0, so.data(), -1.0, Mtmp.data(), tmp.data());
  D.Mpc(tmp.data(), so.data(), Odd, 1);
  const int rc = cg(D, Odd, so.data(), solo.data(), tol, maxit, iters, cg_true_res);
  D.Meooe(solo.data(), tmp.data(), Odd, 0);
  axpby(nh, 1.0, se.data(), -1.0, tmp.data(), Mtmp.data());
  D.MooeeInv(Mtmp.data(), sole.data(), 0);
  for (int64_t s4 = 0; s4 < g.V4; s4++) {
    int x[4];
    g.coor(s4, x);
    const int cb = Geom::parity(x);
    const int64_t h = g.half(x);
    const cd *srcp = (cb == 0 ? sole.data() : solo.data()) + h * Ls * 12;
    std::memcpy(sol + s4 * Ls * 12, srcp, sizeof(cd) * Ls * 12);
  }
  std::vector<cd> r(nf);
  D.M(sol, r.data(), 0);
  axpby(nf, 1.0, r.data(), -1.0, src, r.data());
  *unprec_res = std::sqrt(norm2(nf, r.data()) / norm2(nf, src));
  delete C;
  return rc;
}

}  // extern "C"
