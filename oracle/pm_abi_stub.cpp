// pm_abi_stub.cpp -- a recording stand-in for the C ABI entry points the
// SystemForce adapters call (pm_ctx_create, pm_evaluate, pm_ctx_destroy,
// pm_last_error), so the adapters' translation of ProtoMol's topology and
// force parameters can be checked on a CPU.
//
// TEST INFRASTRUCTURE ONLY.  Each pm_ctx_create appends the force record and
// a summary of the system record to the file $PM_STUB_LOG as key=value
// lines; pm_evaluate adds 1 to every force component and 0.5 to every energy
// slot (add-into semantics are visible to the caller), or, for a vacuum
// system with a region (periodic 0, box > 0) and an atom outside it, adds
// nothing and returns PM_ECAPACITY as the library does.
#include <cstdio>
#include <cstdlib>

#include "protomol_amd.h"

struct pm_ctx {
  int n;
  int region;
  double lo[3], hi[3];
};

static void put_switch(FILE *fp, const char *tag, const pm_switch_desc &s) {
  fprintf(fp, "%s.kind=%d %s.switchon=%.17g %s.cutoff=%.17g %s.switchoff=%.17g %s.order=%.17g ", tag, s.kind, tag,
          s.switchon, tag, s.cutoff, tag, s.switchoff, tag, s.order);
}

extern "C" {

int pm_ctx_create(const pm_force_desc *f, int32_t n_systems, const pm_system_desc *s, pm_ctx **out) {
  const char *path = getenv("PM_STUB_LOG");
  FILE *fp = path ? fopen(path, "a") : 0;
  if (fp) {
    fprintf(fp, "create n_systems=%d term_a=%d term_b=%d cutoff=%.17g flags=%u alpha=%.17g accuracy=%.17g "
                "correction=%d gb_offset=%.17g gb_alpha=%.17g gb_beta=%.17g gb_gamma=%.17g solute=%.17g "
                "solvent=%.17g ace_sigma=%.17g ace_rho_s=%.17g rebuild=%d ",
            n_systems, f->term_a, f->term_b, f->cutoff, f->flags, f->ewald_alpha, f->ewald_accuracy,
            f->ewald_correction, f->gb_offset, f->gb_alpha, f->gb_beta, f->gb_gamma, f->gb_solute_dielec,
            f->gb_solvent_dielec, f->gb_ace_sigma, f->gb_ace_rho_s, f->rebuild_interval);
    put_switch(fp, "sw_a", f->switch_a);
    put_switch(fp, "sw_b", f->switch_b);
    if (n_systems > 0) {
      const pm_system_desc &d = s[0];
      double q = 0.0, m = 0.0, lj = 0.0, gr = 0.0, gs = 0.0;
      long ex = 0;
      for (int i = 0; i < d.n_atoms; i++) {
        q += d.scaled_charge[i] * (i + 1);
        m += d.mass[i];
        if (d.gb_radius) gr += d.gb_radius[i];
        if (d.gb_scaling) gs += d.gb_scaling[i];
      }
      for (int k = 0; k < 4 * d.n_types * d.n_types; k++) lj += d.lj_table[k];
      for (int k = 0; k < d.n_exclusions; k++)
        ex += (long)d.exclusions[3 * k] * 7 + d.exclusions[3 * k + 1] * 3 + d.exclusions[3 * k + 2];
      fprintf(fp, "n_atoms=%d n_types=%d periodic=%d box=%.17g,%.17g,%.17g origin=%.17g,%.17g,%.17g "
                  "charge_w=%.17g mass=%.17g lj_sum=%.17g n_excl=%d excl_w=%ld csf=%.17g gb_r=%.17g gb_s=%.17g",
              d.n_atoms, d.n_types, d.periodic, d.box[0], d.box[1], d.box[2], d.origin[0], d.origin[1],
              d.origin[2], q, m, lj, d.n_exclusions, ex, d.coulomb_scaling, gr, gs);
    }
    fprintf(fp, "\n");
    fclose(fp);
  }
  pm_ctx *c = new pm_ctx();
  c->n = n_systems > 0 ? s[0].n_atoms : 0;
  c->region = n_systems > 0 && !s[0].periodic && s[0].box[0] > 0;
  for (int k = 0; k < 3 && c->region; k++) {
    c->lo[k] = s[0].origin[k] - 0.5 * s[0].box[k];
    c->hi[k] = s[0].origin[k] + 0.5 * s[0].box[k];
  }
  *out = c;
  return PM_OK;
}

int pm_ctx_destroy(pm_ctx *c) {
  delete c;
  return PM_OK;
}

int pm_evaluate(pm_ctx *c, int32_t, const double *x, double *f, double *e) {
  for (int i = 0; i < c->n && c->region; i++)
    for (int k = 0; k < 3; k++)
      if (x[3 * i + k] < c->lo[k] || x[3 * i + k] >= c->hi[k]) return PM_ECAPACITY;
  for (int k = 0; k < 3 * c->n; k++) f[k] += 1.0;
  for (int k = 0; k < PM_E_NSLOTS; k++) e[k] += 0.5;
  return PM_OK;
}

const char *pm_last_error(const pm_ctx *) { return "stub"; }

}  // extern "C"
