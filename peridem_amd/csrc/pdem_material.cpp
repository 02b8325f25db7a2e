// pdem_material.cpp -- per-particle material constants for C++ hosts.
//
// The engine takes each particle's bond-law constants in pdem_particle.mat.
// The reference computes them inside its material classes, whose members are
// private, from the zone's material deck: a C++ DEMModel subclass therefore
// passes the deck values here and gets the same constants.  Formulas and
// operation order follow
//   influence-function moments   src/material/mparticle/influenceFn.cpp:14-122
//   RnpMaterial::computeParameters      material.h:541-620
//   PmbMaterial ctor / computeParameters material.h:674-716, 925-1000
//   PdElastic ctor / computeParameters   material.h:1034-1076, 1260-1306
//   PdState ctor / computeParameters     material.h:1336-1380, 1580-1700
//   inp::MatData conversions             src/inp/decks/materialDeck.h:138-225
// and the bulk modulus of computeMaterialProperties(dim) (material.h:452-489,
// 846-870, 1189-1213, 1504-1530) that BaseParticle turns into Kn.
// Host code only; no device work.
#include <cmath>
#include <string>

#include "pdem.h"

namespace {

constexpr double PI = 3.14159265358979323846;

bool is_greater(double a, double b) {  // util::isGreater (function.cpp:15-23)
  return (a - b) > (std::fabs(a) > std::fabs(b) ? std::fabs(a) : std::fabs(b)) * 1.0E-5;
}
double to_k(double E, double nu) { return E / (3. * (1. - 2. * nu)); }
double to_e(double K, double nu) { return K * (3. * (1. - 2. * nu)); }
double to_lambda_e(double E, double nu) { return E * nu / ((1. + nu) * (1. - 2. * nu)); }
double to_g_e(double E, double nu) { return E / (2. * (1. + nu)); }
double to_e_lambda(double lam, double nu) { return lam * (1. + nu) * (1. - 2. * nu) / nu; }

// i-th moment of the influence function over [0, 1]
bool moment(const pdem_material_deck& d, int i, double* out, std::string* err) {
  const int dim = d.dim;
  const int np = d.n_influence_fn_params;
  const double* p = d.influence_fn_params;
  if (d.influence_fn_type == PDEM_INF_CONST) {
    const double a0 = np > 0 ? p[0] : (double)(dim + 1);
    *out = a0 / (double)(i + 1);
    return true;
  }
  if (d.influence_fn_type == PDEM_INF_LINEAR) {
    double a0, a1;
    if (np == 0) {
      a0 = dim == 1 ? 6. : dim == 2 ? 12. : 20.;
      a1 = -a0;
    } else {
      a0 = p[0];
      a1 = np < 2 ? -a0 : p[1];
    }
    *out = (a0 / (double)(i + 1)) + (a1 / (double)(i + 2));
    return true;
  }
  if (d.influence_fn_type == PDEM_INF_GAUSSIAN) {
    double a, b;
    if (np == 0) {
      b = 0.2;
      if (dim == 1)
        a = 2. / (b * (1. - std::exp(-1. / b)));
      else if (dim == 2)
        a = (4. / b) * 1. / (std::sqrt(PI * b) * std::erf(1. / std::sqrt(b)) - 2. * std::exp(-1. / b));
      else
        a = (2. / b) * 1. / (b - (b + 1.) * std::exp(-1. / b));
    } else {
      a = p[0];
      b = np > 1 ? p[1] : 0.2;
    }
    const double sq1 = std::sqrt(b), sq2 = std::sqrt(PI);
    if (i == 0)
      *out = 0.5 * a * sq1 * sq2 * std::erf(1. / sq1);
    else if (i == 1)
      *out = 0.5 * a * b * (1. - std::exp(-1. / b));
    else if (i == 2)
      *out = 0.5 * a * b * sq1 * (0.5 * sq2 * std::erf(1. / sq1) - (1. / sq1) * std::exp(-1. / b));
    else if (i == 3)
      *out = 0.5 * a * b * b * (1. - (1. + 1. / b) * std::exp(-1. / b));
    else {
      *err = "influence-function moment index must be 0..3";
      return false;
    }
    return true;
  }
  *err = "unknown influence function type";
  return false;
}

// E, nu = 1/4 and lambda of the bond-based laws (E given, or K given)
bool bond_elastic(const pdem_material_deck& d, double* E, double* nu, double* lam, std::string* err) {
  double e = d.E, k = d.K;
  if (e < 0. && k < 0.) {
    *err = "bond-based material needs E or K";
    return false;
  }
  *nu = 0.25;
  if (e > 0.) k = to_k(e, *nu);
  if (k > 0. && e < 0.) e = to_e(k, *nu);
  *E = e;
  *lam = to_lambda_e(e, *nu);
  return true;
}

bool constants(const pdem_material_deck& d, double* mat, double* K_contact, std::string* err) {
  for (int k = 0; k < 6; k++) mat[k] = 0.;
  const double delta = d.horizon;
  if (!(delta > 0.)) {
    *err = "horizon must be positive";
    return false;
  }
  if (d.dim != 2 && d.dim != 3) {
    *err = "dim must be 2 or 3";
    return false;
  }
  switch (d.material_type) {
    case PDEM_MAT_RNP: {
      const double invF = d.dim == 2 ? std::pow(delta, 3) * PI : std::pow(delta, 4) * 4. * PI / 3.;
      double M;
      if (!moment(d, d.dim, &M, err)) return false;
      double C, beta;
      if (d.compute_from_classical) {
        double E = d.E, K = d.K, Gc = d.Gc, KIc = d.KIc;
        if (E < 0. && K < 0.) {
          *err = "RNP needs E or K";
          return false;
        }
        if (Gc < 0. && KIc < 0.) {
          *err = "RNP needs Gc or KIc";
          return false;
        }
        const double nu = 0.25;
        if (E > 0.) K = to_k(E, nu);
        if (K > 0. && E < 0.) E = to_e(K, nu);
        if (Gc > 0.) KIc = std::sqrt(Gc * E);
        if (KIc > 0. && Gc < 0.) Gc = KIc * KIc / E;
        const double lam = to_lambda_e(E, nu);
        if (d.dim == 2) {
          C = PI * Gc / (4. * M);
          beta = 4. * lam / (C * M);
        } else {
          C = 2. * Gc / (3. * M);
          beta = 5. * lam / (C * M);
        }
      } else {
        if (d.n_bond_potential_params < 2) {
          *err = "RNP bond potential needs two parameters";
          return false;
        }
        C = d.bond_potential_params[0];
        beta = d.bond_potential_params[1];
      }
      mat[0] = C;
      mat[1] = beta;
      mat[2] = std::sqrt(0.5 / beta);
      mat[3] = invF;
      mat[4] = d.check_sc_factor;
      const double lam = d.dim == 2 ? C * M * beta / 4. : C * M * beta / 5.;
      *K_contact = to_k(to_e_lambda(lam, 0.25), 0.25);
      return true;
    }
    case PDEM_MAT_PDSTATE: {
      double K, G, s0;
      if (d.compute_from_classical) {
        double E = d.E, nu = d.nu, Gc = d.Gc, KIc = d.KIc;
        K = d.K;
        G = d.G;
        const bool fE = is_greater(E, 0.), fK = is_greater(K, 0.), fG = is_greater(G, 0.),
                   fnu = is_greater(nu, 0.);
        if ((int)fE + (int)fK + (int)fG + (int)fnu != 2) {
          *err = "PDState needs exactly two of E, K, G, nu";
          return false;
        }
        if (!fnu) {
          if (fE && fG) nu = (0.5 * E / G) - 1.;
          if (fE && fK) nu = (3. * K - E) / (6. * K);
          if (fG && fK) nu = (3. * K - 2. * G) / (2. * (3. * K + G));
        }
        if (!fE) {
          if (fK) E = to_e(K, nu);
          if (fG) E = 2. * G * (1. + nu);
        }
        if (!fK) K = to_k(E, nu);
        if (!fG) G = to_g_e(E, nu);
        if (Gc > 0.) KIc = std::sqrt(Gc * E);
        if (KIc > 0. && Gc < 0.) Gc = KIc * KIc / E;
        const double den = (3. * G + std::pow(3. / 4., 4) * (K - 5. * G / 3.)) * delta;
        s0 = std::sqrt(Gc / den);
      } else {
        if (d.n_bond_potential_params < 3) {
          *err = "PDState potential needs three parameters";
          return false;
        }
        K = d.bond_potential_params[0];
        G = d.bond_potential_params[1];
        s0 = d.bond_potential_params[2];
      }
      mat[0] = K;
      mat[1] = G;
      mat[2] = s0;
      *K_contact = K;
      return true;
    }
    case PDEM_MAT_PMB:
    case PDEM_MAT_PDELASTIC: {
      const bool pmb = d.material_type == PDEM_MAT_PMB;
      double c, s0 = 0.;
      if (d.compute_from_classical) {
        double E, nu, lam;
        if (!bond_elastic(d, &E, &nu, &lam, err)) return false;
        if (pmb) {
          double Gc = d.Gc;
          const double KIc = d.KIc;
          if (Gc < 0. && KIc < 0.) {
            *err = "PMB needs Gc or KIc";
            return false;
          }
          if (KIc > 0. && Gc < 0.) Gc = KIc * KIc / E;
          c = d.dim == 2 ? 24. * E / (PI * std::pow(delta, 3.) * (1. - nu))
                         : 24. * lam / (PI * std::pow(delta, 3.));
          s0 = std::sqrt(5. * PI * Gc / (9. * E * delta));
        } else {
          c = 24. * lam / (PI * std::pow(delta, 3.));
        }
      } else {
        if (d.n_bond_potential_params < (pmb ? 2 : 1)) {
          *err = "bond potential parameters missing";
          return false;
        }
        c = d.bond_potential_params[0];
        if (pmb) s0 = d.bond_potential_params[1];
      }
      mat[0] = c;
      mat[1] = s0;
      const double lam_p = c * (PI * std::pow(delta, 3.)) / 24.;
      *K_contact = to_k(to_e_lambda(lam_p, 0.25), 0.25);
      return true;
    }
    default:
      *err = "material type not supported (RNPBond, PMBBond, PDElasticBond, PDState)";
      return false;
  }
}

thread_local std::string g_material_err;

}  // namespace

extern "C" {

int pdem_material_constants(const pdem_material_deck* deck, double* mat, double* K_contact) {
  if (!deck || !mat || !K_contact) {
    g_material_err = "null argument";
    return -1;
  }
  std::string err;
  if (!constants(*deck, mat, K_contact, &err)) {
    g_material_err = err;
    return -1;
  }
  g_material_err.clear();
  return 0;
}

const char* pdem_material_last_error(void) { return g_material_err.c_str(); }

}  // extern "C"
