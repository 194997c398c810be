/*
 * oracle.c -- CPU restatement of the sjavis/minim hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry and bench.py's cpu_baseline leg load it (liboracle.so); the product
 * library (minim_amd/) never links, loads or calls it, and has no CPU fallback.
 *
 * Serial FP64, written to follow the reference's operation order so that with
 * -ffp-contract=off it agrees bit for bit with a single-thread build of the
 * reference (oracle/_ref, built from the reference's own sources by
 * oracle/Makefile).
 *
 * Parity pins (see tests/test_oracle.py):
 *   orc_pf_*            reference unit-test known answers (PhaseField_test.cpp:31-68,
 *                       130-145), bit-exact vs oracle/_ref on random grids, finite
 *                       differences, symmetry E(-phi)=E(phi), translation invariance.
 *   orc_lbfgs_minimise  bit-exact trajectory vs oracle/_ref (iterations, energy,
 *                       coordinates); n-step convergence on quadratics.
 *   orc_fire_minimise   Fire_test.cpp:25-40 exact first step; convergence KAT
 *                       Fire_test.cpp:43-52; bit-exact vs oracle/_ref.
 *   orc_fire_minimise_ex  options without a reference counterpart: off, it is
 *                       orc_fire_minimise bit for bit; dt0 gives the closed-form
 *                       first step -0.5 dt0^2 g; the max-force criterion holds at exit.
 *   orc_graddescent_*   bit-exact vs oracle/_ref.
 *   orc_lj_*            State_test.cpp:33-51 dimer KAT; vs oracle/_ref; Newton's 3rd law.
 *   orc_ljpbc_*         NO reference implementation exists (the reference LjNd has
 *                       no cutoff or box).  Pinned by: equality with orc_lj_* when
 *                       rc exceeds every pair distance and the box is large; E
 *                       continuous and zero at r = rc; finite differences; sum g = 0.
 *   orc_bah_*           BarAndHinge_test.cpp:13-56 minima; vs oracle/_ref; the
 *                       reference's 2x stretching-gradient factor (BarAndHinge.cpp:282)
 *                       checked against a finite difference of the energy.
 *   orc_pfx_*           PhaseField's remaining terms: PhaseField_test.cpp:44-197 known
 *                       answers; vs oracle/_ref on 40 random layouts (energies,
 *                       constrained gradients, projections, energyComponents, node
 *                       volumes and surface areas; 1e-13), L-BFGS trajectories and
 *                       diffuseSolid vs oracle/_ref; finite differences; closed forms of
 *                       pressure and confinement; zero volume change and idempotence of
 *                       the volume projection (tests/test_pf_ext_oracle.py).
 *   orc_pf_site_types   site for site vs the reference build's getType, including the
 *                       layouts it rejects (tests/test_pf_ext_oracle.py).
 *   orc_pfu_*           PhaseFieldUnstructured_test.cpp:27-80,104-129,132-180 known
 *                       answers; vs oracle/_ref on 40 random layouts (element counts,
 *                       energies, gradients, projections; 1e-13) and L-BFGS
 *                       trajectories; finite differences (tests/test_pfu_oracle.py).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* PhaseField                                                                */
/* ------------------------------------------------------------------------- */

/* PhaseField::assignFluidCoefficients (src/potentials/PhaseField.cpp:78-134),
 * PhaseField::init (:137-158), setDefaults (:961-967), checkArraySizes (:970-995).
 * Only MODEL_BASIC with spatially uniform coefficients is restated. */
int orc_pf_init(orc_pf* pf, const int grid[3], int nfluid, double resolution,
                const double* surface_tension, int n_st,
                const double* interface_size, int n_iw,
                int density, double density_const, const unsigned char* fix_fluid) {
  memset(pf, 0, sizeof(*pf));
  if (nfluid < 1 || nfluid > ORC_MAXF) return 1;
  for (int d = 0; d < 3; d++) {
    if (grid[d] < 1) return 1;
    pf->grid[d] = grid[d];
  }
  pf->nfluid = nfluid;
  pf->resolution = resolution;
  pf->density = density;
  pf->density_const = density_const;
  int nparams = nfluid; /* PhaseField.cpp:141, MODEL_BASIC */

  double st[ORC_MAXF], iw[ORC_MAXF];
  if (n_st == 1) {
    for (int i = 0; i < nparams; i++) st[i] = surface_tension[0];
  } else if (n_st == nparams) {
    for (int i = 0; i < nparams; i++) st[i] = surface_tension[i];
  } else {
    return 2; /* "Invalid size of surfaceTension array." :86 */
  }
  /* vec::sum is std::accumulate from 0 in order, then / size (:88) */
  double sum = 0;
  for (int i = 0; i < nparams; i++) sum += st[i];
  pf->st_mean = sum / nparams;

  /* setDefaults: interfaceSize := resolution for every fluid if unset (:962) */
  if (n_iw == 0) {
    for (int i = 0; i < nfluid; i++) iw[i] = resolution;
    n_iw = nfluid;
  } else if (n_iw == 1) {
    for (int i = 0; i < nparams; i++) iw[i] = interface_size[0];
  } else if (n_iw == nparams) {
    for (int i = 0; i < nparams; i++) iw[i] = interface_size[i];
  } else {
    return 3; /* "Invalid size of interfaceSize array." :95 */
  }

  if (nfluid <= 2) {
    /* kappa = 3*st/iw, kappaP = 3*st*iw  (:110-111) */
    for (int i = 0; i < nparams; i++) {
      pf->kappa[i] = (3 * st[i]) / iw[i];
      pf->kappaP[i] = (3 * st[i]) * iw[i];
    }
  } else {
    /* ternary combination (:116-131) */
    double ks[ORC_MAXF], kps[ORC_MAXF];
    for (int i = 0; i < nparams; i++) {
      ks[i] = (6 * st[i]) / iw[i];
      kps[i] = (6 * st[i]) * iw[i];
    }
    int l = nfluid - 1;
    pf->kappa[0] = 0.5 * (ks[0] + ks[1] - ks[l]);
    pf->kappa[1] = 0.5 * (ks[0] - ks[1] + ks[l]);
    pf->kappa[2] = 0.5 * (-ks[0] + ks[1] + ks[l]);
    pf->kappaP[0] = 0.5 * (kps[0] + kps[1] - kps[l]);
    pf->kappaP[1] = 0.5 * (kps[0] - kps[1] + kps[l]);
    pf->kappaP[2] = 0.5 * (-kps[0] + kps[1] + kps[l]);
    for (int i = 3; i < nfluid; i++) {
      pf->kappa[i] = ks[i - 1] - pf->kappa[0];
      pf->kappaP[i] = kps[i - 1] - pf->kappaP[0];
    }
  }
  pf->convergence = 1e-8 * pf->st_mean * pow(resolution, 2); /* :146 */
  for (int i = 0; i < nfluid; i++) pf->fix[i] = fix_fluid ? fix_fluid[i] : 0;
  if (density == ORC_DENSITY_FIXED) pf->fix[nfluid - 1] = 1; /* :147 */
  return 0;
}

static int pf_idx(const orc_pf* pf, int x, int y, int z) {
  /* getIdx with periodic wrap (PhaseField.cpp:44-49) */
  const int* n = pf->grid;
  x = (x + n[0]) % n[0];
  y = (y + n[1]) % n[1];
  z = (z + n[2]) % n[2];
  return (x * n[1] + y) * n[2] + z;
}

/* PhaseField::energyGradient (PhaseField.cpp:662-703) with one thread, no solid
 * nodes, nodeVol = resolution^3 everywhere (:194-208 with type 0), calling
 * fluidEnergy (:365-393) -> phaseGradient (:238-288, "no solid" branch :252-263)
 * and densityConstraintEnergy (:457-473).  Pressure, surface, force and
 * confinement terms are inactive (zero) in this restatement.
 * Caller semantics of State::basicEG (src/State.cpp:48-56): e is set (not
 * accumulated) and g is zeroed before the scatter. */
void orc_pf_energy_gradient(const orc_pf* pf, const double* x, double* e, double* g) {
  const int nF = pf->nfluid;
  const int* n = pf->grid;
  const size_t ngrid = (size_t)n[0] * n[1] * n[2];
  const double nodeVol = 1 * pow(pf->resolution, 3);
  double eLocal = 0;
  if (g) memset(g, 0, sizeof(double) * ngrid * nF);

  for (int ix = 0; ix < n[0]; ix++) {
    for (int iy = 0; iy < n[1]; iy++) {
      for (int iz = 0; iz < n[2]; iz++) {
        int iGrid = pf_idx(pf, ix, iy, iz);
        int nb[6];
        nb[0] = pf_idx(pf, ix - 1, iy, iz);
        nb[1] = pf_idx(pf, ix + 1, iy, iz);
        nb[2] = pf_idx(pf, ix, iy - 1, iz);
        nb[3] = pf_idx(pf, ix, iy + 1, iz);
        nb[4] = pf_idx(pf, ix, iy, iz - 1);
        nb[5] = pf_idx(pf, ix, iy, iz + 1);
        /* fluidEnergy */
        for (int iF = 0; iF < nF; iF++) {
          int iDof = iGrid * nF + iF;
          double c = x[iDof];
          if (nF == 1) {
            double factor = pf->kappa[iF] / 16 * nodeVol;
            if (e) eLocal += factor * pow(c + 1, 2) * pow(c - 1, 2);
            if (g) g[iDof] += factor * 4 * c * (c * c - 1);
          } else {
            double factor = 0.5 * pf->kappa[iF] * nodeVol;
            if (e) eLocal += factor * pow(c, 2) * pow(c - 1, 2);
            if (g) g[iDof] += factor * 2 * c * (c - 1) * (2 * c - 1);
          }
          double factor = (nF == 1) ? 0.25 * pf->kappaP[iF] * nodeVol : 0.5 * pf->kappaP[iF] * nodeVol;
          factor = factor / pow(pf->resolution, 2);
          /* phaseGradient */
          double grad2 = 0;
          for (int iDir = 0; iDir < 3; iDir++) {
            if (n[iDir] == 1) continue;
            int im = nb[2 * iDir] * nF + iF;
            int ip = nb[2 * iDir + 1] * nF + iF;
            double gradm = c - x[im];
            double gradp = c - x[ip];
            if (e) grad2 += 0.5 * (pow(gradm, 2) + pow(gradp, 2));
            if (g) {
              g[iDof] += factor * (gradm + gradp);
              g[im] -= factor * gradm;
              g[ip] -= factor * gradp;
            }
          }
          if (e) eLocal += factor * grad2;
        }
        /* densityConstraintEnergy (soft constraint) */
        if (nF > 1 && pf->density == ORC_DENSITY_SOFT) {
          double coef = pf->density_const * pf->st_mean * pow(pf->resolution, 2);
          double rhoDiff = -1;
          for (int iF = 0; iF < nF; iF++) rhoDiff += x[iGrid * nF + iF];
          if (e) eLocal += coef * pow(rhoDiff, 2);
          if (g) {
            for (int iF = 0; iF < nF; iF++) g[iGrid * nF + iF] += 2 * coef * rhoDiff;
          }
        }
      }
    }
  }
  if (e) {
    double tot = 0;
    tot += eLocal; /* basicEG zeroes *e, the OpenMP region adds eLocal (:697-701) */
    *e = tot;
  }
}

/* PhaseField::specialisedConstraints (PhaseField.cpp:580-612): zero fixed
 * fluids, then remove the per-site mean over the variable fluids (hard density
 * constraint).  The volume projection (:616-658) is inactive here. */
void orc_pf_constraints(const orc_pf* pf, double* data) {
  const int nF = pf->nfluid;
  const size_t ngrid = (size_t)pf->grid[0] * pf->grid[1] * pf->grid[2];
  int var[ORC_MAXF], nvar = 0;
  for (int iF = 0; iF < nF; iF++) {
    if (pf->fix[iF]) {
      for (size_t i = 0; i < ngrid; i++) data[i * nF + iF] = 0;
    } else {
      var[nvar++] = iF;
    }
  }
  if (nF > 1 && pf->density == ORC_DENSITY_HARD) {
    double normFactor = 1.0 / nvar;
    for (size_t i = 0; i < ngrid; i++) {
      double component = 0;
      for (int k = 0; k < nvar; k++) component += data[i * nF + var[k]];
      component *= normFactor;
      for (int k = 0; k < nvar; k++) data[i * nF + var[k]] -= component;
    }
  }
}

static void pf_eg_cb(const void* ctx, const double* x, double* e, double* g) {
  orc_pf_energy_gradient((const orc_pf*)ctx, x, e, g);
}
static void pf_con_cb(const void* ctx, const double* x, double* data) {
  (void)x;
  orc_pf_constraints((const orc_pf*)ctx, data);
}
void orc_pf_potential(const orc_pf* pf, orc_pot* out) {
  out->ctx = pf;
  out->ndof = (size_t)pf->grid[0] * pf->grid[1] * pf->grid[2] * pf->nfluid;
  out->eg = pf_eg_cb;
  out->constrain = pf_con_cb;
  out->convergence = pf->convergence;
}

/* ------------------------------------------------------------------------- */
/* PhaseField: solid nodes, wetting, pressure, force, confinement, volume    */
/* ------------------------------------------------------------------------- */

/* PhaseField::getType (src/potentials/PhaseField.cpp:899-958).  The 26
 * neighbours are counted by how many of their offsets are non-zero: faces (1),
 * edges (2) and corners (3); "one side" (type 4 against 5) asks whether the
 * solid face's axis has no solid edge in the plane across it (the reference's
 * neighbour groups 10-13, 7/8/15/16 and 6/9/14/17).  Offsets wrap periodically
 * (getIdx, :44-49), so a dimension of size 1 or 2 can count one site twice, as
 * in the reference. */
int orc_pf_site_types(const int grid[3], const unsigned char* solid, int* type) {
  const int n0 = grid[0], n1 = grid[1], n2 = grid[2];
  for (int x = 0; x < n0; x++)
    for (int y = 0; y < n1; y++)
      for (int z = 0; z < n2; z++) {
        const int i = (x * n1 + y) * n2 + z;
        if (solid[i]) {
          type[i] = -1;
          continue;
        }
        int nface = 0, nedge = 0, ncorner = 0;
        int face_axis[3] = {0, 0, 0};     /* a solid face along axis a */
        int edge_across[3] = {0, 0, 0};   /* a solid edge with zero offset along axis a */
        for (int dx = -1; dx <= 1; dx++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
              const int nz = (dx != 0) + (dy != 0) + (dz != 0);
              if (nz == 0) continue;
              const int j = (((x + dx + n0) % n0) * n1 + (y + dy + n1) % n1) * n2 + (z + dz + n2) % n2;
              if (!solid[j]) continue;
              if (nz == 1) {
                nface++;
                face_axis[dx != 0 ? 0 : (dy != 0 ? 1 : 2)] = 1;
              } else if (nz == 2) {
                nedge++;
                if (dx == 0) edge_across[0] = 1;
                if (dy == 0) edge_across[1] = 1;
                if (dz == 0) edge_across[2] = 1;
              } else {
                ncorner++;
              }
            }
        int t = -2;
        if (nface + nedge + ncorner == 0) {
          t = 0;
        } else if (nface == 0) {
          if (nedge == 0 && ncorner == 1) t = 1;
          else if (nedge == 1 && ncorner <= 2) t = 2;
          else if (nedge == 2 && ncorner <= 3) t = 3;
        } else if (nface == 1) {
          int one = 0;
          for (int a = 0; a < 3; a++)
            if (face_axis[a] && !edge_across[a]) one = 1;
          t = one ? 4 : 5;
        } else if (nface == 2) {
          t = 6;
        } else if (nface == 3) {
          t = 7;
        }
        if (t == -2) return 1; /* "Undefined surface type" (:957) */
        type[i] = t;
      }
  return 0;
}

void orc_pfx_free(orc_pfx* p) {
  free(p->kappa);
  free(p->kappaP);
  free(p->solid);
  free(p->type);
  free(p->node_vol);
  free(p->surface_area);
  free(p->contact_angle);
  free(p->ff_init);
  memset(p, 0, sizeof(*p));
}

/* assignFluidCoefficients (:78-134) with per-site tensions, init (:137-158),
 * initLocal (:161-235) on one rank, setDefaults (:961-967), checkArraySizes
 * (:970-995). */
int orc_pfx_init(orc_pfx* p, const int grid[3], int nfluid, double resolution,
                 const double* surface_tension, int n_st, const double* interface_size, int n_iw,
                 int density, double density_const, const unsigned char* fix_fluid,
                 const unsigned char* solid, const double* contact_angle, int n_ca,
                 const double* pressure, const double* force, const double* confinement,
                 int volume_fixed, double vol_const, const double* volume, const double* x0) {
  memset(p, 0, sizeof(*p));
  if (nfluid < 1 || nfluid > ORC_MAXF) return 1;
  for (int d = 0; d < 3; d++)
    if (grid[d] < 1) return 1;
  const size_t ng = (size_t)grid[0] * grid[1] * grid[2];
  const int nF = nfluid, np = nfluid;
  p->ngrid = ng;
  orc_pf* b = &p->pf;
  for (int d = 0; d < 3; d++) b->grid[d] = grid[d];
  b->nfluid = nF;
  b->resolution = resolution;
  b->density = density;
  b->density_const = density_const;
  const int nk = (n_st == (int)(ng * np)) ? (int)ng : 1; /* :79 */
  p->nkgrid = nk;
  /* checkArraySizes: interfaceSize has nfluid entries after setDefaults (the
   * scalar setter, n_iw == 1, fills nfluid copies) */
  double iw1[ORC_MAXF];
  if (n_iw == 0) {
    for (int i = 0; i < nF; i++) iw1[i] = resolution;
  } else if (n_iw == 1) {
    for (int i = 0; i < nF; i++) iw1[i] = interface_size[0];
  } else if (n_iw == nF) {
    for (int i = 0; i < nF; i++) iw1[i] = interface_size[i];
  } else {
    return 3;
  }
  if (n_st != 1 && n_st != nk * np) return 2; /* :84-87 */
  const size_t nst = (size_t)nk * np;
  double* st = (double*)malloc(sizeof(double) * nst);
  for (size_t i = 0; i < nst; i++) st[i] = (n_st == 1) ? surface_tension[0] : surface_tension[i];
  double sum = 0;
  for (size_t i = 0; i < nst; i++) sum += st[i];
  b->st_mean = sum / nst; /* :88 */
  p->kappa = (double*)malloc(sizeof(double) * nst);
  p->kappaP = (double*)malloc(sizeof(double) * nst);
  for (int ig = 0; ig < nk; ig++) {
    const double* s = st + (size_t)ig * np;
    double* ka = p->kappa + (size_t)ig * np;
    double* kp = p->kappaP + (size_t)ig * np;
    if (nF <= 2) {
      for (int i = 0; i < np; i++) {
        ka[i] = (3 * s[i]) / iw1[i];
        kp[i] = (3 * s[i]) * iw1[i];
      }
    } else {
      double ks[ORC_MAXF], kps[ORC_MAXF];
      for (int i = 0; i < np; i++) {
        ks[i] = (6 * s[i]) / iw1[i];
        kps[i] = (6 * s[i]) * iw1[i];
      }
      const int l = nF - 1;
      ka[0] = 0.5 * (ks[0] + ks[1] - ks[l]);
      ka[1] = 0.5 * (ks[0] - ks[1] + ks[l]);
      ka[2] = 0.5 * (-ks[0] + ks[1] + ks[l]);
      kp[0] = 0.5 * (kps[0] + kps[1] - kps[l]);
      kp[1] = 0.5 * (kps[0] - kps[1] + kps[l]);
      kp[2] = 0.5 * (-kps[0] + kps[1] + kps[l]);
      for (int i = 3; i < nF; i++) {
        ka[i] = ks[i - 1] - ka[0];
        kp[i] = kps[i - 1] - kp[0];
      }
    }
  }
  free(st);
  for (int i = 0; i < nF && i < ORC_MAXF; i++) {
    b->kappa[i] = p->kappa[i];
    b->kappaP[i] = p->kappaP[i];
  }
  b->convergence = 1e-8 * b->st_mean * pow(resolution, 2); /* :146 */
  for (int i = 0; i < nF; i++) b->fix[i] = fix_fluid ? fix_fluid[i] : 0;
  if (density == ORC_DENSITY_FIXED) b->fix[nF - 1] = 1;
  if (contact_angle && n_ca != (int)(ng * nF)) {
    orc_pfx_free(p);
    return 4; /* "Invalid size of contactAngle array." :977 */
  }

  /* forces (:149-157): magnitude and unit direction per fluid */
  if (force) {
    for (int i = 0; i < nF; i++) {
      const double* f = force + 3 * i;
      double r = 0;
      for (int d = 0; d < 3; d++) r += f[d] * f[d];
      p->fmag[i] = sqrt(r);
      for (int d = 0; d < 3; d++) p->fnorm[i][d] = f[d] / p->fmag[i];
      if (p->fmag[i] != 0) p->any_force = 1; /* vec::any(fMag) */
    }
  }
  for (int i = 0; i < nF; i++) {
    p->pressure[i] = pressure ? pressure[i] : 0;
    p->conf[i] = confinement ? confinement[i] : 0;
  }

  p->solid = (unsigned char*)calloc(ng, 1);
  if (solid) memcpy(p->solid, solid, ng);
  p->type = (int*)malloc(sizeof(int) * ng);
  if (orc_pf_site_types(grid, p->solid, p->type)) {
    orc_pfx_free(p);
    return 5;
  }
  /* node volume and solid surface area per site (:194-208) */
  p->node_vol = (double*)malloc(sizeof(double) * ng);
  p->surface_area = (double*)malloc(sizeof(double) * ng);
  for (size_t i = 0; i < ng; i++) {
    const int t = p->type[i];
    double v = 0, a = 0;
    if (t == 0) {
      v = 1;
    } else if (t > 0) {
      v = t / 8.0;
      if (t == 2 || t == 4 || t == 6) a = 1;
      if (t == 1 || t == 7) a = 0.75;
      if (t == 3 || t == 5) a = 1.25;
    }
    p->node_vol[i] = v * pow(resolution, 3);
    p->surface_area[i] = a * pow(resolution, 2);
  }
  if (contact_angle) {
    p->contact_angle = (double*)malloc(sizeof(double) * n_ca);
    memcpy(p->contact_angle, contact_angle, sizeof(double) * n_ca);
  }

  /* fixed-volume target and sum of squared node volumes (:216-231) */
  p->volume_fixed = volume_fixed;
  p->vol_const = vol_const;
  if (volume) {
    p->has_volume = 1;
    for (int i = 0; i < nF; i++) p->volume[i] = volume[i];
  } else if (volume_fixed) {
    p->has_volume = 1;
    for (int i = 0; i < nF; i++) p->volume[i] = 0;
    for (size_t ig = 0; ig < ng; ig++)
      for (int i = 0; i < nF; i++) p->volume[i] += x0[i + nF * ig] * p->node_vol[ig];
  }
  /* the reference sets totalVolume2 only when it derives the target itself;
   * with an explicit target it is left unset there.  It is always set here. */
  p->total_volume2 = 0;
  for (size_t ig = 0; ig < ng; ig++) p->total_volume2 += p->node_vol[ig] * p->node_vol[ig];

  int any_conf = 0;
  for (int i = 0; i < nF; i++)
    if (p->conf[i] != 0) any_conf = 1;
  if (any_conf) {
    p->ff_init = (double*)malloc(sizeof(double) * ng * nF);
    memcpy(p->ff_init, x0, sizeof(double) * ng * nF);
  }
  return 0;
}

/* The terms of one site, each adding to its own energy/gradient pair (the
 * reference's per-site calls in energyGradient, :679-694, and
 * energyComponents, :720-729).  t[c] selects term c. */
enum { PFX_FLUID, PFX_SURFACE, PFX_PRESSURE, PFX_DENSITY, PFX_FORCE, PFX_CONF, PFX_NTERM };

static void pfx_site(const orc_pfx* p, const double* x, int ix, int iy, int iz, double** e, double** g) {
  const orc_pf* b = &p->pf;
  const int nF = b->nfluid;
  const int* n = b->grid;
  const double res = b->resolution;
  const int iGrid = pf_idx(b, ix, iy, iz);
  const double v = p->node_vol[iGrid];
  int nb[6];
  nb[0] = pf_idx(b, ix - 1, iy, iz);
  nb[1] = pf_idx(b, ix + 1, iy, iz);
  nb[2] = pf_idx(b, ix, iy - 1, iz);
  nb[3] = pf_idx(b, ix, iy + 1, iz);
  nb[4] = pf_idx(b, ix, iy, iz - 1);
  nb[5] = pf_idx(b, ix, iy, iz + 1);

  /* fluidEnergy (:365-393) with phaseGradient (:238-288) */
  if (e[PFX_FLUID] || g[PFX_FLUID]) {
    double* E = e[PFX_FLUID];
    double* G = g[PFX_FLUID];
    for (int iF = 0; iF < nF; iF++) {
      const int iDof = iGrid * nF + iF;
      const double c = x[iDof];
      const int iK = (p->nkgrid == 1) ? iF : iGrid * nF + iF;
      if (nF == 1) {
        double factor = p->kappa[iK] / 16 * v;
        if (E) *E += factor * pow(c + 1, 2) * pow(c - 1, 2);
        if (G) G[iDof] += factor * 4 * c * (c * c - 1);
      } else {
        double factor = 0.5 * p->kappa[iK] * v;
        if (E) *E += factor * pow(c, 2) * pow(c - 1, 2);
        if (G) G[iDof] += factor * 2 * c * (c - 1) * (2 * c - 1);
      }
      double factor = (nF == 1) ? 0.25 * p->kappaP[iK] * v : 0.5 * p->kappaP[iK] * v;
      factor = factor / pow(res, 2);
      double grad2 = 0;
      for (int d = 0; d < 3; d++) {
        if (n[d] == 1) continue;
        const int mg = nb[2 * d], pg = nb[2 * d + 1];
        const int im = mg * nF + iF, ip = pg * nF + iF;
        if (!p->solid[mg] && !p->solid[pg]) {
          const double gm = c - x[im], gp = c - x[ip];
          if (E) grad2 += 0.5 * (pow(gm, 2) + pow(gp, 2));
          if (G) {
            G[iDof] += factor * (gm + gp);
            G[im] -= factor * gm;
            G[ip] -= factor * gp;
          }
        } else if (!p->solid[pg]) {
          const double gp = c - x[ip];
          if (E) grad2 += pow(gp, 2);
          if (G) {
            G[iDof] += factor * 2 * gp;
            G[ip] -= factor * 2 * gp;
          }
        } else if (!p->solid[mg]) {
          const double gm = c - x[im];
          if (E) grad2 += pow(gm, 2);
          if (G) {
            G[iDof] += factor * 2 * gm;
            G[im] -= factor * 2 * gm;
          }
        }
      }
      if (E) *E += factor * grad2;
    }
  }

  /* surfaceEnergy (:476-487): binary fluids only */
  if ((e[PFX_SURFACE] || g[PFX_SURFACE]) && p->surface_area[iGrid] != 0 && p->contact_angle &&
      p->contact_angle[iGrid] != 90 && nF == 1) {
    const double w = 1 / sqrt(2.0) * cos(p->contact_angle[iGrid] * 3.1415926536 / 180);
    const double phi = x[iGrid];
    if (e[PFX_SURFACE]) *e[PFX_SURFACE] += w * (pow(phi, 3) / 3 - phi - 2.0 / 3) * p->surface_area[iGrid];
    if (g[PFX_SURFACE]) g[PFX_SURFACE][iGrid] += w * (pow(phi, 2) - 1) * p->surface_area[iGrid];
  }

  /* pressureEnergy (:396-421) */
  if (e[PFX_PRESSURE] || g[PFX_PRESSURE]) {
    double* E = e[PFX_PRESSURE];
    double* G = g[PFX_PRESSURE];
    for (int iF = 0; iF < nF; iF++) {
      if (p->pressure[iF] == 0 || b->fix[iF]) continue;
      if (nF == 1) {
        const double vol = 0.5 * (x[iGrid] + 1) * v;
        if (E) *E -= p->pressure[iF] * vol;
        if (G) G[iGrid] -= 0.5 * p->pressure[iF] * v;
      } else {
        const int iDof = iGrid * nF + iF;
        const double vol = x[iDof] * v;
        if (E) *E -= p->pressure[iF] * vol;
        if (G) G[iDof] -= p->pressure[iF] * v;
      }
    }
  }

  /* densityConstraintEnergy (:424-440) */
  if ((e[PFX_DENSITY] || g[PFX_DENSITY]) && nF > 1 && b->density == ORC_DENSITY_SOFT) {
    const double coef = b->density_const * b->st_mean * pow(res, 2);
    double rho = -1;
    for (int iF = 0; iF < nF; iF++) rho += x[iGrid * nF + iF];
    if (e[PFX_DENSITY]) *e[PFX_DENSITY] += coef * pow(rho, 2);
    if (g[PFX_DENSITY])
      for (int iF = 0; iF < nF; iF++) g[PFX_DENSITY][iGrid * nF + iF] += 2 * coef * rho;
  }

  /* forceEnergy (:490-518).  Without a gradient the reference returns after
   * the first fluid that contributes (:511), so an energy-only call counts
   * one fluid; this is kept. */
  if ((e[PFX_FORCE] || g[PFX_FORCE]) && !p->solid[iGrid] && p->any_force) {
    const double coord[3] = {ix - (n[0] - 1) / 2.0, iy - (n[1] - 1) / 2.0, iz - (n[2] - 1) / 2.0};
    for (int iF = 0; iF < nF; iF++) {
      if (p->fmag[iF] == 0) continue;
      const int iDof = iGrid * nF + iF;
      const double c = (nF == 1) ? 0.5 * (1 + x[iDof]) : x[iDof];
      if (c < 0.01) continue;
      double dot = 0;
      for (int d = 0; d < 3; d++) dot += coord[d] * p->fnorm[iF][d];
      const double h = -dot * res;
      const double ePerConc = v * p->fmag[iF] * h;
      if (e[PFX_FORCE]) *e[PFX_FORCE] += c * ePerConc;
      if (!g[PFX_FORCE]) break;
      if (nF == 1) g[PFX_FORCE][iDof] += 0.5 * ePerConc;
      else g[PFX_FORCE][iDof] += ePerConc;
    }
  }

  /* ffConfinementEnergy (:521-537) */
  if (e[PFX_CONF] || g[PFX_CONF]) {
    for (int iF = 0; iF < nF; iF++) {
      if (p->conf[iF] == 0) continue;
      const int iDof = iGrid * nF + iF;
      const double coef = p->conf[iF] * b->st_mean * pow(res, 2);
      const double c = x[iDof], c0 = p->ff_init[iDof];
      if ((c0 - 0.5) * (c - 0.5) < 0) {
        if (e[PFX_CONF]) *e[PFX_CONF] += coef * pow(c - 0.5, 2);
        if (g[PFX_CONF]) g[PFX_CONF][iDof] += coef * 2 * (c - 0.5);
      }
    }
  }
}

/* PhaseField::energyGradient (:662-703) on one thread: every term adds to one
 * energy and one gradient; State::basicEG zeroes both first. */
void orc_pfx_energy_gradient(const orc_pfx* p, const double* x, double* e, double* g) {
  const int* n = p->pf.grid;
  const size_t nd = p->ngrid * p->pf.nfluid;
  double eLocal = 0;
  double* E[PFX_NTERM];
  double* G[PFX_NTERM];
  for (int t = 0; t < PFX_NTERM; t++) {
    E[t] = e ? &eLocal : NULL;
    G[t] = g;
  }
  if (g) memset(g, 0, sizeof(double) * nd);
  for (int ix = 0; ix < n[0]; ix++)
    for (int iy = 0; iy < n[1]; iy++)
      for (int iz = 0; iz < n[2]; iz++) pfx_site(p, x, ix, iy, iz, E, G);
  if (e) *e = 0 + eLocal;
}

/* specialisedConstraints (:580-659): fixed fluids, hard density projection,
 * then removal of the volume-changing component of every variable fluid. */
void orc_pfx_constraints(const orc_pfx* p, double* data) {
  orc_pf_constraints(&p->pf, data);
  if (!p->volume_fixed) return;
  const int nF = p->pf.nfluid;
  int var[ORC_MAXF], nvar = 0;
  for (int iF = 0; iF < nF; iF++)
    if (!p->pf.fix[iF]) var[nvar++] = iF;
  const int hard = p->pf.density == ORC_DENSITY_HARD;
  double comp[ORC_MAXF] = {0};
  for (size_t ig = 0; ig < p->ngrid; ig++) {
    const double v = p->node_vol[ig];
    for (int a = 0; a < nvar; a++) {
      comp[var[a]] += data[ig * nF + var[a]] * v;
      if (hard)
        for (int b2 = 0; b2 < nvar; b2++) comp[var[a]] -= data[ig * nF + var[b2]] * v / nvar;
    }
  }
  for (int a = 0; a < nvar; a++) comp[var[a]] = comp[var[a]] / p->total_volume2;
  for (size_t ig = 0; ig < p->ngrid; ig++) {
    const double v = p->node_vol[ig];
    for (int a = 0; a < nvar; a++) {
      data[ig * nF + var[a]] -= comp[var[a]] * v;
      if (hard)
        for (int b2 = 0; b2 < nvar; b2++) data[ig * nF + var[b2]] += comp[var[a]] * v / nvar;
    }
  }
}

/* PhaseField::energyComponents (:706-744) with volumeConstraintEnergy
 * (:540-577) on one rank. */
void orc_pfx_components(const orc_pfx* p, const double* x, double out[14]) {
  const int* n = p->pf.grid;
  const int nF = p->pf.nfluid;
  const size_t nd = p->ngrid * nF;
  double ev[PFX_NTERM + 1] = {0};
  double* gv[PFX_NTERM + 1];
  for (int t = 0; t <= PFX_NTERM; t++) gv[t] = (double*)calloc(nd, sizeof(double));
  double* E[PFX_NTERM];
  for (int t = 0; t < PFX_NTERM; t++) E[t] = &ev[t];
  for (int ix = 0; ix < n[0]; ix++)
    for (int iy = 0; iy < n[1]; iy++)
      for (int iz = 0; iz < n[2]; iz++) pfx_site(p, x, ix, iy, iz, E, gv);
  if (p->volume_fixed) {
    double vol[ORC_MAXF] = {0}, diff[ORC_MAXF] = {0};
    for (size_t ig = 0; ig < p->ngrid; ig++) {
      if (nF == 1) {
        vol[0] += 0.5 * (x[ig] + 1) * p->node_vol[ig];
      } else {
        for (int iF = 0; iF < nF; iF++) vol[iF] += x[ig * nF + iF] * p->node_vol[ig];
      }
    }
    for (int iF = 0; iF < nF; iF++) {
      if (p->pf.fix[iF] || p->volume[iF] < 0) continue;
      diff[iF] = vol[iF] - p->volume[iF];
    }
    const double coef = p->vol_const * p->pf.st_mean / pow(p->pf.resolution, 4);
    double* g = gv[PFX_NTERM];
    for (int iF = 0; iF < nF; iF++) {
      if (diff[iF] == 0) continue;
      ev[PFX_NTERM] += coef * pow(diff[iF], 2) / 1;
      for (size_t ig = 0; ig < p->ngrid; ig++) {
        const size_t iDof = ig * nF + iF;
        double w = 4 * x[iDof] * (1 - x[iDof]);
        if (w < 0.0) w = 0.0;
        g[iDof] += coef * diff[iF] * p->node_vol[ig] * w;
      }
    }
  }
  for (int t = 0; t <= PFX_NTERM; t++) {
    double s = 0;
    for (size_t i = 0; i < nd; i++) s += gv[t][i] * gv[t][i];
    out[2 * t] = ev[t];
    out[2 * t + 1] = sqrt(s);
    free(gv[t]);
  }
}

static void pfx_eg_cb(const void* ctx, const double* x, double* e, double* g) {
  orc_pfx_energy_gradient((const orc_pfx*)ctx, x, e, g);
}
static void pfx_con_cb(const void* ctx, const double* x, double* data) {
  (void)x;
  orc_pfx_constraints((const orc_pfx*)ctx, data);
}
void orc_pfx_potential(const orc_pfx* p, orc_pot* out) {
  out->ctx = p;
  out->ndof = p->ngrid * p->pf.nfluid;
  out->eg = pfx_eg_cb;
  out->constrain = pfx_con_cb;
  out->convergence = p->pf.convergence;
}

/* ------------------------------------------------------------------------- */
/* PhaseFieldUnstructured: the element form of the phase field               */
/* ------------------------------------------------------------------------- */

enum { PFU_FLUID = 0, PFU_FLUID_ALL = 1, PFU_PRESSURE = 3, PFU_DENSITY = 4, PFU_SURFACE = 5, PFU_FORCE = 6,
       PFU_CONF = 7 };

void orc_pfu_free(orc_pfu* p) {
  free(p->node_vol);
  free(p->el_type);
  free(p->el_idof);
  free(p->el_par);
  memset(p, 0, sizeof(*p));
}

static void pfu_push(orc_pfu* p, size_t* cap, int type, const int* idof, int n, const double* par, int npar) {
  if (p->nel == *cap) {
    *cap = *cap ? 2 * *cap : 1024;
    p->el_type = (int*)realloc(p->el_type, sizeof(int) * *cap);
    p->el_idof = (int*)realloc(p->el_idof, sizeof(int) * ORC_PFU_NIDOF * *cap);
    p->el_par = (double*)realloc(p->el_par, sizeof(double) * 3 * *cap);
  }
  p->el_type[p->nel] = type;
  for (int i = 0; i < ORC_PFU_NIDOF; i++) p->el_idof[p->nel * ORC_PFU_NIDOF + i] = i < n ? idof[i] : -1;
  for (int i = 0; i < 3; i++) p->el_par[p->nel * 3 + i] = i < npar ? par[i] : 0;
  p->nel++;
}

/* PhaseFieldUnstructured::assignFluidCoefficients (PhaseFieldUnstructured.cpp:71-107),
 * init (:110-257), initLocal on one rank (:260-277), setDefaults, checkArraySizes. */
int orc_pfu_init(orc_pfu* p, const int grid[3], int nfluid, double resolution, const double* surface_tension,
                 int n_st, const double* interface_size, int n_iw, int density, double density_const,
                 const unsigned char* fix_fluid, const unsigned char* solid, const double* contact_angle, int n_ca,
                 const double* pressure, const double* force, const double* confinement, int volume_fixed,
                 double vol_const, const double* volume, const double* x0) {
  memset(p, 0, sizeof(*p));
  if (nfluid < 1 || nfluid > ORC_MAXF) return 1;
  for (int d = 0; d < 3; d++) {
    if (grid[d] < 1) return 1;
    p->grid[d] = grid[d];
  }
  const int nF = nfluid;
  const size_t ng = (size_t)grid[0] * grid[1] * grid[2];
  p->ngrid = ng;
  p->nfluid = nF;
  p->resolution = resolution;
  p->density = density;
  p->density_const = density_const;
  /* the soft constraint's element reads three DOFs of each site
     (PhaseFieldUnstructured.cpp:635): with two fluids that reads past the element */
  if (nF == 2 && density == ORC_DENSITY_SOFT) return 6;
  double iw[ORC_MAXF], st[ORC_MAXF];
  if (n_iw == 0) {
    for (int i = 0; i < nF; i++) iw[i] = resolution;
  } else if (n_iw == 1) {
    for (int i = 0; i < nF; i++) iw[i] = interface_size[0];
  } else if (n_iw == nF) {
    for (int i = 0; i < nF; i++) iw[i] = interface_size[i];
  } else {
    return 3;
  }
  if (n_st == 1) {
    for (int i = 0; i < nF; i++) st[i] = surface_tension[0];
  } else if (n_st == nF) {
    for (int i = 0; i < nF; i++) st[i] = surface_tension[i];
  } else {
    return 2;
  }
  double sum = 0;
  for (int i = 0; i < nF; i++) sum += st[i];
  p->st_mean = sum / nF;
  if (nF <= 2) {
    for (int i = 0; i < nF; i++) {
      p->kappa[i] = 3 * st[i] / iw[i];
      p->kappaP[i] = 3 * st[i] * iw[i];
    }
  } else {
    double ks[ORC_MAXF], kps[ORC_MAXF];
    for (int i = 0; i < nF; i++) {
      ks[i] = 6 * st[i] / iw[i];
      kps[i] = 6 * st[i] * iw[i];
    }
    const int l = nF - 1;
    p->kappa[0] = 0.5 * (ks[0] + ks[1] - ks[l]);
    p->kappa[1] = 0.5 * (ks[0] - ks[1] + ks[l]);
    p->kappa[2] = 0.5 * (-ks[0] + ks[1] + ks[l]);
    p->kappaP[0] = 0.5 * (kps[0] + kps[1] - kps[l]);
    p->kappaP[1] = 0.5 * (kps[0] - kps[1] + kps[l]);
    p->kappaP[2] = 0.5 * (-kps[0] + kps[1] + kps[l]);
    for (int i = 3; i < nF; i++) { /* :103-104, unlike PhaseField.cpp:128-129 */
      p->kappa[i] = 0.5 * (-ks[0] - ks[1] + ks[l]) + ks[i - 1];
      p->kappaP[i] = 0.5 * (-kps[0] - kps[1] + kps[l]) + kps[i - 1];
    }
  }
  p->convergence = 1e-8 * p->st_mean * pow(resolution, 2);
  for (int i = 0; i < nF; i++) p->fix[i] = fix_fluid ? fix_fluid[i] : 0;
  if (density == ORC_DENSITY_FIXED) p->fix[nF - 1] = 1;
  if (contact_angle && n_ca != (int)(ng * nF)) return 4;

  double fmag[ORC_MAXF] = {0}, fnorm[ORC_MAXF][3];
  if (force) {
    for (int i = 0; i < nF; i++) {
      double r = 0;
      for (int d = 0; d < 3; d++) r += force[3 * i + d] * force[3 * i + d];
      fmag[i] = sqrt(r);
      for (int d = 0; d < 3; d++) fnorm[i][d] = force[3 * i + d] / fmag[i];
    }
  }
  unsigned char* sol = (unsigned char*)calloc(ng, 1);
  if (solid) memcpy(sol, solid, ng);
  int* type = (int*)malloc(sizeof(int) * ng);
  if (orc_pf_site_types(grid, sol, type)) {
    free(sol);
    free(type);
    return 5;
  }
  p->node_vol = (double*)malloc(sizeof(double) * ng);
  double* area = (double*)malloc(sizeof(double) * ng);
  for (size_t i = 0; i < ng; i++) {
    const int t = type[i];
    double v = 0, a = 0;
    if (t == 0) {
      v = 1;
    } else if (t > 0) {
      v = t / 8.0;
      if (t == 2 || t == 4 || t == 6) a = 1;
      if (t == 1 || t == 7) a = 0.75;
      if (t == 3 || t == 5) a = 1.25;
    }
    p->node_vol[i] = v * pow(resolution, 3);
    area[i] = a * pow(resolution, 2);
  }
  p->volume_fixed = volume_fixed || volume != NULL;
  p->vol_const = vol_const;
  if (volume) {
    for (int i = 0; i < nF; i++) p->volume[i] = volume[i];
  } else if (p->volume_fixed) {
    for (size_t ig = 0; ig < ng; ig++)
      for (int i = 0; i < nF; i++) p->volume[i] += x0[i + nF * ig] * p->node_vol[ig];
  }

  /* elements, site by site (:157-240) */
  size_t cap = 0;
  const int n0 = grid[0], n1 = grid[1], n2 = grid[2];
  static const int off[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}};
  for (int x = 0; x < n0; x++)
    for (int y = 0; y < n1; y++)
      for (int z = 0; z < n2; z++) {
        const int iGrid = (x * n1 + y) * n2 + z;
        if (sol[iGrid]) continue;
        int nodes[7];
        nodes[0] = iGrid;
        for (int k = 0; k < 6; k++) {
          const int j = (((x + off[k][0] + n0) % n0) * n1 + (y + off[k][1] + n1) % n1) * n2 + (z + off[k][2] + n2) % n2;
          nodes[k + 1] = sol[j] ? iGrid : j;
        }
        const double vol = p->node_vol[iGrid];
        if (density == ORC_DENSITY_FIXED) {
          int idof[7];
          for (int k = 0; k < 7; k++) idof[k] = nodes[k] * nF;
          pfu_push(p, &cap, PFU_FLUID_ALL, idof, 7, &vol, 1);
        } else {
          for (int f = 0; f < nF; f++) {
            int idof[7];
            for (int k = 0; k < 7; k++) idof[k] = nodes[k] * nF + f;
            const double par[2] = {vol, (double)f};
            pfu_push(p, &cap, PFU_FLUID, idof, 7, par, 2);
          }
        }
        for (int f = 0; f < nF; f++) {
          if (!pressure || pressure[f] == 0) continue;
          const int idof = iGrid * nF + f;
          const double par = pressure[f];
          pfu_push(p, &cap, PFU_PRESSURE, &idof, 1, &par, 1);
        }
        if (nF > 1 && density == ORC_DENSITY_SOFT) {
          int idof[ORC_PFU_NIDOF];
          const int n = nF < ORC_PFU_NIDOF ? nF : ORC_PFU_NIDOF;
          for (int f = 0; f < n; f++) idof[f] = iGrid * nF + f;
          pfu_push(p, &cap, PFU_DENSITY, idof, n, NULL, 0);
        }
        if (area[iGrid] > 0 && contact_angle && contact_angle[iGrid] != 90) {
          const double w = 1 / sqrt(2.0) * cos(contact_angle[iGrid] * 3.1415926536 / 180);
          for (int f = 0; f < nF; f++) {
            const int idof = iGrid * nF + f;
            const double par[2] = {area[iGrid], w};
            pfu_push(p, &cap, PFU_SURFACE, &idof, 1, par, 2);
          }
        }
        for (int f = 0; f < nF; f++) {
          if (!(fmag[f] > 0)) continue;
          const double coord[3] = {x - (n0 - 1) / 2.0, y - (n1 - 1) / 2.0, z - (n2 - 1) / 2.0};
          double dot = 0;
          for (int d = 0; d < 3; d++) dot += coord[d] * fnorm[f][d];
          const double h = -dot * resolution;
          const double par[3] = {vol, fmag[f], h};
          const int idof = iGrid * nF + f;
          pfu_push(p, &cap, PFU_FORCE, &idof, 1, par, 3);
        }
        for (int f = 0; f < nF; f++) {
          if (!confinement || confinement[f] == 0) continue;
          const int idof = iGrid * nF + f;
          const double par[2] = {confinement[f], x0[idof]};
          pfu_push(p, &cap, PFU_CONF, &idof, 1, par, 2);
        }
      }
  free(sol);
  free(type);
  free(area);
  return 0;
}

/* phaseGradient of the element form (PhaseFieldUnstructured.cpp:311-515): a
 * neighbour replaced by the site itself marks a solid one. */
static void pfu_phase_gradient(const double* x, const int* idof, double factor, double* e, double* g) {
  static const int mp[3][2] = {{1, 6}, {2, 5}, {3, 4}};
  const double c1 = x[idof[0]];
  double grad2 = 0;
  for (int d = 0; d < 3; d++) {
    const int im = idof[mp[d][0]], ip = idof[mp[d][1]];
    if (im != idof[0] && ip != idof[0]) {
      const double gradm = c1 - x[im], gradp = c1 - x[ip];
      if (e) grad2 += 0.5 * (pow(gradm, 2) + pow(gradp, 2));
      if (g) {
        g[idof[0]] += factor * (gradm + gradp);
        g[im] -= factor * gradm;
        g[ip] -= factor * gradp;
      }
    } else if (ip != idof[0]) {
      const double gradp = c1 - x[ip];
      if (e) grad2 += pow(gradp, 2);
      if (g) {
        g[idof[0]] += factor * 2 * gradp;
        g[ip] -= factor * 2 * gradp;
      }
    } else if (im != idof[0]) {
      const double gradm = c1 - x[im];
      if (e) grad2 += pow(gradm, 2);
      if (g) {
        g[idof[0]] += factor * 2 * gradm;
        g[im] -= factor * 2 * gradm;
      }
    }
  }
  if (e) *e += factor * grad2;
}

/* elementEnergyGradient (:695-724) of one element */
static void pfu_element(const orc_pfu* p, const double* x, size_t ie, double* e, double* g) {
  const int nF = p->nfluid;
  const int* idof = p->el_idof + ie * ORC_PFU_NIDOF;
  const double* par = p->el_par + ie * 3;
  const double res = p->resolution;
  switch (p->el_type[ie]) {
    case PFU_FLUID: { /* fluidEnergy (:518-538) */
      const double vol = par[0];
      const int f = (int)par[1];
      const double c = x[idof[0]];
      if (nF == 1) {
        const double factor = p->kappa[0] / 16 * vol;
        if (e) *e += factor * pow(c + 1, 2) * pow(c - 1, 2);
        if (g) g[idof[0]] += factor * 4 * c * (c * c - 1);
      } else {
        const double factor = 0.5 * p->kappa[f] * vol;
        if (e) *e += factor * pow(c, 2) * pow(c - 1, 2);
        if (g) g[idof[0]] += factor * 2 * c * (c - 1) * (2 * c - 1);
      }
      double factor = (nF == 1) ? 0.25 * p->kappaP[0] * vol : 0.5 * p->kappaP[f] * vol;
      factor = factor / pow(res, 2);
      pfu_phase_gradient(x, idof, factor, e, g);
      break;
    }
    case PFU_FLUID_ALL: { /* fluidEnergyAll (:541-586) */
      const double vol = par[0];
      double cN[7], gN[7] = {0};
      static const int iota[7] = {0, 1, 2, 3, 4, 5, 6};
      for (int i = 0; i < 7; i++) cN[i] = 1;
      for (int f = 0; f < nF - 1; f++)
        for (int i = 0; i < 7; i++) cN[i] -= x[idof[i] + f];
      const double bfN = 0.5 * p->kappa[nF - 1] * vol;
      const double gfN = 0.5 * p->kappaP[nF - 1] * vol; /* not divided by res^2 (:555) */
      if (e) *e += bfN * pow(cN[0], 2) * pow(cN[0] - 1, 2);
      if (g) {
        gN[0] += bfN * 2 * cN[0] * (cN[0] - 1) * (2 * cN[0] - 1);
        pfu_phase_gradient(cN, iota, gfN, e, gN);
      } else {
        pfu_phase_gradient(cN, iota, gfN, e, NULL);
      }
      for (int f = 0; f < nF - 1; f++) {
        int id[7];
        for (int i = 0; i < 7; i++) id[i] = idof[i] + f;
        const double c = x[id[0]];
        const double bf = 0.5 * p->kappa[f] * vol;
        if (e) *e += bf * pow(c, 2) * pow(c - 1, 2);
        if (g) g[id[0]] += bf * 2 * c * (c - 1) * (2 * c - 1);
        double gf = 0.5 * p->kappaP[f] * vol;
        gf = gf / pow(res, 2);
        pfu_phase_gradient(x, id, gf, e, g);
        if (g)
          for (int i = 0; i < 7; i++) g[id[i]] -= gN[i];
      }
      break;
    }
    case PFU_PRESSURE: { /* pressureEnergy (:613-628) */
      const int iDof = idof[0];
      const double pr = par[0];
      const double v = p->node_vol[iDof / nF];
      if (nF == 1) {
        const double volume = 0.5 * (x[iDof] + 1) * v;
        if (e) *e -= pr * volume;
        if (g) g[iDof] -= 0.5 * pr * v;
      } else {
        const double volume = x[iDof] * v;
        if (e) *e -= pr * volume;
        if (g) g[iDof] -= pr * v;
      }
      break;
    }
    case PFU_DENSITY: { /* densityConstraintEnergy (:631-642): three DOFs */
      if (nF == 1) break;
      const double coef = p->density_const * p->st_mean * pow(res, 2);
      const double rho = x[idof[0]] + x[idof[1]] + x[idof[2]] - 1;
      if (e) *e += coef * pow(rho, 2);
      if (g) {
        g[idof[0]] += 2 * coef * rho;
        g[idof[1]] += 2 * coef * rho;
        g[idof[2]] += 2 * coef * rho;
      }
      break;
    }
    case PFU_SURFACE: /* surfaceEnergy (:645-654) */
      if (nF == 1) {
        const double phi = x[idof[0]];
        if (e) *e += par[1] * (pow(phi, 3) / 3 - phi - 2.0 / 3) * par[0];
        if (g) g[idof[0]] += par[1] * (pow(phi, 2) - 1) * par[0];
      }
      break;
    case PFU_FORCE: { /* forceEnergy (:657-676): every fluid, also without a gradient */
      const double c = (nF == 1) ? 0.5 * (1 + x[idof[0]]) : x[idof[0]];
      if (c < 0.01) break;
      const double vol = par[0], f = par[1], h = par[2];
      if (e) *e += c * vol * f * h;
      if (g) {
        if (nF == 1) g[idof[0]] += 0.5 * vol * f * h;
        else g[idof[0]] += vol * f * h;
      }
      break;
    }
    case PFU_CONF: { /* ffConfinementEnergy (:679-692) */
      const double c = x[idof[0]];
      const double coef = par[0] * p->st_mean * pow(res, 2);
      const double c0 = par[1];
      if ((c0 - 0.5) * (c - 0.5) < 0) {
        if (e) *e += coef * pow(c - 0.5, 2);
        if (g) g[idof[0]] += coef * 2 * (c - 0.5);
      }
      break;
    }
  }
}

/* State's elementEG (src/State.cpp:58-79) without the constraints: e set, g
 * zeroed, every element in order, then blockEnergyGradient (:280-308). */
void orc_pfu_energy_gradient(const orc_pfu* p, const double* x, double* e, double* g) {
  const int nF = p->nfluid;
  const size_t nd = p->ngrid * nF;
  if (e) *e = 0;
  if (g) memset(g, 0, sizeof(double) * nd);
  for (size_t ie = 0; ie < p->nel; ie++) pfu_element(p, x, ie, e, g);
  if (!p->volume_fixed) return;
  double vol[ORC_MAXF] = {0};
  for (size_t iDof = 0; iDof < nd; iDof++) {
    const double v = p->node_vol[iDof / nF];
    if (nF == 1) vol[0] += 0.5 * (x[iDof] + 1) * v;
    else vol[iDof % nF] += x[iDof] * v;
  }
  const double coef = p->vol_const * p->st_mean / pow(p->resolution, 4);
  for (int f = 0; f < nF; f++) {
    const double diff = vol[f] - p->volume[f];
    if (e) *e += coef * pow(diff, 2) / 1;
    if (!g) continue;
    for (size_t ig = 0; ig < p->ngrid; ig++) {
      const size_t iDof = ig * nF + f;
      const double w0 = 4 * x[iDof] * (1 - x[iDof]);
      const double w = w0 > 0.0 ? w0 : 0.0;
      g[iDof] += coef * diff * p->node_vol[ig] * w;
    }
  }
}

/* Potential::applyConstraints (src/Potential.cpp:102-115) with the constraints
 * PhaseFieldUnstructured::init sets (:242-256): per site the variable fluids'
 * sum (hard density constraint), then the fixed fluids' DOFs. */
void orc_pfu_constraints(const orc_pfu* p, double* data) {
  const int nF = p->nfluid;
  int var[ORC_MAXF], nvar = 0;
  for (int f = 0; f < nF; f++)
    if (!p->fix[f]) var[nvar++] = f;
  if (nF > 1 && p->density == ORC_DENSITY_HARD) {
    for (size_t ig = 0; ig < p->ngrid; ig++) {
      double nn = 0, gn = 0;
      for (int a = 0; a < nvar; a++) nn += 1.0 * 1.0;
      for (int a = 0; a < nvar; a++) gn += data[ig * nF + var[a]] * 1.0;
      for (int a = 0; a < nvar; a++) data[ig * nF + var[a]] -= gn * 1.0 / nn;
    }
  }
  for (int f = 0; f < nF; f++) {
    if (!p->fix[f]) continue;
    for (size_t ig = 0; ig < p->ngrid; ig++) data[ig * nF + f] = 0;
  }
}

static void pfu_eg_cb(const void* ctx, const double* x, double* e, double* g) {
  orc_pfu_energy_gradient((const orc_pfu*)ctx, x, e, g);
}
static void pfu_con_cb(const void* ctx, const double* x, double* data) {
  (void)x;
  orc_pfu_constraints((const orc_pfu*)ctx, data);
}
void orc_pfu_potential(const orc_pfu* p, orc_pot* out) {
  out->ctx = p;
  out->ndof = p->ngrid * p->nfluid;
  out->eg = pfu_eg_cb;
  out->constrain = pfu_con_cb;
  out->convergence = p->convergence;
}

/* ------------------------------------------------------------------------- */
/* State-level helpers                                                       */
/* ------------------------------------------------------------------------- */

/* CommGrid::dotProduct on one rank (src/communicators/CommGrid.cpp:187-194):
 * serial loop over RangeI = every index in order. */
double orc_dot(const double* a, const double* b, size_t n) {
  double v = 0;
  for (size_t i = 0; i < n; i++) v += a[i] * b[i];
  return v;
}

typedef struct {
  const orc_pot* pot;
  long ne, ng;
} orc_state;

/* State::energy(coords) (src/State.cpp:134-152) -> basicEG with e only. */
static double st_energy(orc_state* s, const double* x) {
  double e = 0;
  s->ne++;
  s->pot->eg(s->pot->ctx, x, &e, NULL);
  return e;
}

/* State::procGradient (src/State.cpp:243-248) -> basicEG with g, then
 * Potential::applyConstraints (src/Potential.cpp:102-115). */
static void st_gradient(orc_state* s, const double* x, double* g) {
  s->ng++;
  s->pot->eg(s->pot->ctx, x, NULL, g);
  if (s->pot->constrain) s->pot->constrain(s->pot->ctx, x, g);
}

static void st_constrain(orc_state* s, const double* x, double* data) {
  if (s->pot->constrain) s->pot->constrain(s->pot->ctx, x, data);
}

/* backtrackingLinesearch (src/minimisers/linesearch.cpp:8-29).  step is
 * updated in place; x becomes x+step (the last, possibly unevaluated, trial). */
static double linesearch(orc_state* s, double* x, double* step, double de0, double* tmp, size_t n) {
  const double c = 0.5, tau = 0.5;
  double t = -c * de0;
  double e0 = st_energy(s, x);
  double mult = 1;
  for (size_t j = 0; j < n; j++) tmp[j] = x[j] + step[j];
  for (int i = 0; i < 10; i++) {
    double e = st_energy(s, tmp);
    if (e0 - e >= t) break;
    for (size_t j = 0; j < n; j++) step[j] = step[j] * tau;
    for (size_t j = 0; j < n; j++) tmp[j] = x[j] + step[j];
    mult = mult * tau;
    t = t * tau;
  }
  memcpy(x, tmp, sizeof(double) * n);
  return mult;
}

/* ------------------------------------------------------------------------- */
/* L-BFGS                                                                    */
/* ------------------------------------------------------------------------- */

/* Minimiser::minimise (src/Minimiser.cpp:27-38) driving Lbfgs::init
 * (src/minimisers/Lbfgs.cpp:34-38), Lbfgs::iteration (:41-81),
 * Lbfgs::getDirection (:84-127) and Lbfgs::checkConvergence (:130-133). */
int orc_lbfgs_minimise(const orc_pot* pot, double* x, int m, int max_iter, double max_step,
                       double init_step, int ls, double convergence, orc_stats* st) {
  size_t n = pot->ndof;
  orc_state s = {pot, 0, 0};
  double* g = calloc(n, sizeof(double));
  double* gNew = calloc(n, sizeof(double));
  double* step = calloc(n, sizeof(double));
  double* tmp = calloc(n, sizeof(double));
  double* S = calloc((size_t)m * n, sizeof(double));
  double* Y = calloc((size_t)m * n, sizeof(double));
  double* rho = calloc(m, sizeof(double));
  double* alpha = calloc(m, sizeof(double));
  if (!g || !gNew || !step || !tmp || !S || !Y || !rho || !alpha) return 1;
  int I = 0, iter, conv = 0;
  for (iter = 0; iter <= max_iter; iter++) {
    /* ---- Lbfgs::iteration ---- */
    if (iter == 0) {
      st_gradient(&s, x, g);
      I = 0;
    } else {
      I++;
    }
    /* getDirection */
    int m_tmp = I < m ? I : m;
    int ic = I % m;
    for (size_t j = 0; j < n; j++) step[j] = -g[j];
    if (I == 0) {
      double gnorm = sqrt(orc_dot(g, g, n));
      if (gnorm > 0) {
        double f = init_step / gnorm;
        for (size_t j = 0; j < n; j++) step[j] *= f;
      }
    } else {
      for (int i1 = 0; i1 < m_tmp; i1++) {
        int k = ((ic - 1 - i1) + m) % m;
        alpha[k] = rho[k] * orc_dot(step, S + (size_t)k * n, n);
        const double* yk = Y + (size_t)k * n;
        for (size_t j = 0; j < n; j++) step[j] -= alpha[k] * yk[j];
      }
      int k = (ic - 1 + m) % m;
      double gamma = 1 / (rho[k] * orc_dot(Y + (size_t)k * n, Y + (size_t)k * n, n));
      for (size_t j = 0; j < n; j++) step[j] *= gamma;
      for (int i1 = 0; i1 < m_tmp; i1++) {
        int kk = ((ic - m_tmp + i1) + m) % m;
        double beta = rho[kk] * orc_dot(step, Y + (size_t)kk * n, n);
        double f = alpha[kk] - beta;
        const double* sk = S + (size_t)kk * n;
        for (size_t j = 0; j < n; j++) {
          double t = f * sk[j];
          step[j] += t;
        }
      }
      if (max_step != 0) {
        double ss = sqrt(orc_dot(step, step, n));
        if (ss > max_step) {
          double f = max_step / ss;
          for (size_t j = 0; j < n; j++) step[j] *= f;
        }
      }
    }
    st_constrain(&s, x, step);
    double gs = orc_dot(g, step, n);
    if (gs > 0) {
      gs = -gs;
      for (size_t j = 0; j < n; j++) step[j] = -step[j];
    }
    if (ls == ORC_LS_BACKTRACKING) {
      linesearch(&s, x, step, gs, tmp, n);
    } else {
      for (size_t j = 0; j < n; j++) x[j] = x[j] + step[j];
    }
    st_gradient(&s, x, gNew);
    for (size_t j = 0; j < n; j++) tmp[j] = gNew[j] - g[j];
    double sy = orc_dot(step, tmp, n);
    if (sy != 0) {
      int k = I % m;
      memcpy(S + (size_t)k * n, step, sizeof(double) * n);
      memcpy(Y + (size_t)k * n, tmp, sizeof(double) * n);
      rho[k] = 1 / sy;
    } else {
      I--;
    }
    memcpy(g, gNew, sizeof(double) * n);
    /* checkConvergence */
    double rms = sqrt(orc_dot(g, g, n) / n);
    if (rms < convergence) {
      conv = 1;
      break;
    }
  }
  if (st) {
    st->iterations = iter;
    st->converged = conv;
    st->grad_rms = sqrt(orc_dot(g, g, n) / n);
    double e = 0;
    pot->eg(pot->ctx, x, &e, NULL);
    st->energy = e;
    st->n_energy = s.ne;
    st->n_gradient = s.ng;
  }
  free(g); free(gNew); free(step); free(tmp); free(S); free(Y); free(rho); free(alpha);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* FIRE                                                                      */
/* ------------------------------------------------------------------------- */

/* Fire::init (src/minimisers/Fire.cpp:27-38), Fire::iteration (:41-84),
 * Fire::checkConvergence (:87-90); constants include/minimisers/Fire.h:24-28.
 * dt_max == 0 requests the reference's estimate 0.1/sqrt(|g|). */
/* Largest Euclidean norm over groups of dim consecutive entries (the largest
 * per-particle force when g is a gradient of dim-dimensional particles). */
static double group_norm_max(const double* g, size_t n, int dim) {
  double m = 0;
  for (size_t q = 0; q + (size_t)dim <= n; q += (size_t)dim) {
    double s = 0;
    for (int k = 0; k < dim; k++) s += g[q + k] * g[q + k];
    if (s > m || s != s) m = s;
  }
  return sqrt(m);
}

int orc_fire_minimise(const orc_pot* pot, double* x, double dt_max, int max_iter,
                      int ls, double convergence, orc_stats* st, double* dt_max_out) {
  return orc_fire_minimise_ex(pot, x, dt_max, 0.0, 0, max_iter, ls, convergence, st, dt_max_out);
}

/* Fire with two options the reference does not have (both off reproduce it):
 * dt0 > 0 starts from that time step instead of dtMax (Fire.cpp:43), and
 * max_force_dim > 0 converges on the largest per-particle force norm over
 * groups of that many DOFs instead of the RMS gradient (Fire.cpp:87-90). */
int orc_fire_minimise_ex(const orc_pot* pot, double* x, double dt_max, double dt0, int max_force_dim,
                         int max_iter, int ls, double convergence, orc_stats* st, double* dt_max_out) {
  const int nMin = 5;
  const double fInc = 1.1, fDec = 0.5, fA = 0.99, aStart = 0.1;
  size_t n = pot->ndof;
  orc_state s = {pot, 0, 0};
  double* g = calloc(n, sizeof(double));
  double* v = calloc(n, sizeof(double));
  double* step = calloc(n, sizeof(double));
  double* tmp = calloc(n, sizeof(double));
  if (!g || !v || !step || !tmp) return 1;
  double gNorm = 0, a = aStart, dt = 0;
  int nSteps = 0;
  if (dt_max == 0) {
    st_gradient(&s, x, g);
    gNorm = sqrt(orc_dot(g, g, n));
    dt_max = (gNorm != 0) ? 0.1 / sqrt(gNorm) : 1;
  }
  int iter, conv = 0;
  for (iter = 0; iter <= max_iter; iter++) {
    if (iter == 0) {
      dt = dt0 > 0 ? dt0 : dt_max;
      st_gradient(&s, x, g);
      gNorm = sqrt(orc_dot(g, g, n));
    }
    double p = -orc_dot(v, g, n);
    if (p > 0) {
      double vNorm = sqrt(orc_dot(v, v, n));
      double f1 = 1 - a;
      double f2 = a * vNorm / gNorm + dt;
      for (size_t j = 0; j < n; j++) {
        double t1 = f1 * v[j];
        double t2 = f2 * g[j];
        v[j] = t1 - t2;
      }
      nSteps++;
    } else {
      double f = -dt;
      for (size_t j = 0; j < n; j++) v[j] = f * g[j];
      nSteps = 0;
    }
    if ((p > 0) && (nSteps > nMin)) {
      a = a * fA;
      dt = dt * fInc;
      if (dt > dt_max) dt = dt_max;
    } else if (p <= 0) {
      a = aStart;
      dt = dt * fDec;
    }
    for (size_t j = 0; j < n; j++) step[j] = dt * v[j];
    st_constrain(&s, x, step);
    if (ls == ORC_LS_BACKTRACKING) {
      double gs = orc_dot(g, step, n);
      linesearch(&s, x, step, gs, tmp, n);
    } else {
      for (size_t j = 0; j < n; j++) x[j] = x[j] + step[j];
    }
    st_gradient(&s, x, g);
    gNorm = sqrt(orc_dot(g, g, n));
    const double crit = max_force_dim > 0 ? group_norm_max(g, n, max_force_dim) : gNorm / sqrt((double)n);
    if (crit < convergence) {
      conv = 1;
      break;
    }
  }
  if (st) {
    st->iterations = iter;
    st->converged = conv;
    st->grad_rms = gNorm / sqrt((double)n);
    double e = 0;
    pot->eg(pot->ctx, x, &e, NULL);
    st->energy = e;
    st->n_energy = s.ne;
    st->n_gradient = s.ng;
  }
  if (dt_max_out) *dt_max_out = dt_max;
  free(g); free(v); free(step); free(tmp);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Gradient descent                                                          */
/* ------------------------------------------------------------------------- */

/* GradDescent::iteration / checkConvergence (src/minimisers/GradDescent.cpp:27-46). */
int orc_graddescent_minimise(const orc_pot* pot, double* x, double alpha, int max_iter,
                             int ls, double convergence, orc_stats* st) {
  size_t n = pot->ndof;
  orc_state s = {pot, 0, 0};
  double* g = calloc(n, sizeof(double));
  double* step = calloc(n, sizeof(double));
  double* tmp = calloc(n, sizeof(double));
  if (!g || !step || !tmp) return 1;
  int iter, conv = 0;
  for (iter = 0; iter <= max_iter; iter++) {
    st_gradient(&s, x, g);
    double f = -alpha;
    for (size_t j = 0; j < n; j++) step[j] = f * g[j];
    if (ls == ORC_LS_BACKTRACKING) {
      double gs = orc_dot(g, step, n);
      linesearch(&s, x, step, gs, tmp, n);
    } else {
      for (size_t j = 0; j < n; j++) x[j] = x[j] + step[j];
    }
    if (sqrt(orc_dot(g, g, n) / n) < convergence) {
      conv = 1;
      break;
    }
  }
  if (st) {
    st->iterations = iter;
    st->converged = conv;
    st->grad_rms = sqrt(orc_dot(g, g, n) / n);
    double e = 0;
    pot->eg(pot->ctx, x, &e, NULL);
    st->energy = e;
    st->n_energy = s.ne;
    st->n_gradient = s.ng;
  }
  free(g); free(step); free(tmp);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Lennard-Jones                                                             */
/* ------------------------------------------------------------------------- */

/* LjNd::init element order (src/potentials/LjNd.cpp:12-26: all i<j pairs) and
 * LjNd::elementEnergyGradient (:29-53), accumulated as State::elementEG
 * (src/State.cpp:58-79). */
void orc_lj_energy_gradient(const orc_lj* lj, const double* x, double* e, double* g) {
  const int D = lj->ndim, N = lj->nparticle;
  double etot = 0;
  if (g) memset(g, 0, sizeof(double) * (size_t)N * D);
  double dx[8];
  for (int i = 0; i < N; i++) {
    for (int j = i + 1; j < N; j++) {
      double r2 = 0;
      for (int d = 0; d < D; d++) {
        dx[d] = x[D * i + d] - x[D * j + d];
        r2 += dx[d] * dx[d];
      }
      double lj6 = pow(lj->sigma, 6) / pow(r2, 3);
      double lj12 = lj6 * lj6;
      if (e) etot += 4 * lj->epsilon * (lj12 - lj6);
      if (g) {
        double factor = -24 * lj->epsilon * (2 * lj12 - lj6) / r2;
        for (int d = 0; d < D; d++) {
          g[D * i + d] += dx[d] * factor;
          g[D * j + d] -= dx[d] * factor;
        }
      }
    }
  }
  if (e) *e = etot;
}

static void lj_eg_cb(const void* ctx, const double* x, double* e, double* g) {
  orc_lj_energy_gradient((const orc_lj*)ctx, x, e, g);
}
void orc_lj_potential(const orc_lj* lj, size_t ndof, orc_pot* out) {
  out->ctx = lj;
  out->ndof = ndof;
  out->eg = lj_eg_cb;
  out->constrain = NULL;
  out->convergence = 1e-6; /* Potential::convergence default (include/Potential.h:25) */
}

/* Truncated-and-shifted LJ in a periodic cube, minimum image, all pairs.
 * E_pair = 4 eps [(s/r)^12 - (s/r)^6] - 4 eps [(s/rc)^12 - (s/rc)^6] for r < rc.
 * New semantics (the reference has no cutoff or periodic box); pinned in
 * tests against orc_lj_energy_gradient for rc > all distances, continuity at
 * rc, finite differences and momentum conservation. */
void orc_ljpbc_energy_gradient(const orc_ljpbc* lj, const double* x, double* e, double* g) {
  const int N = lj->nparticle;
  const double L = lj->box, rc2 = lj->rc * lj->rc;
  double src6 = pow(lj->sigma / lj->rc, 6);
  double shift = 4 * lj->epsilon * (src6 * src6 - src6);
  double etot = 0;
  if (g) memset(g, 0, sizeof(double) * (size_t)N * 3);
  for (int i = 0; i < N; i++) {
    for (int j = i + 1; j < N; j++) {
      double dx[3], r2 = 0;
      for (int d = 0; d < 3; d++) {
        double t = x[3 * i + d] - x[3 * j + d];
        t -= L * nearbyint(t / L);
        dx[d] = t;
        r2 += t * t;
      }
      if (r2 >= rc2) continue;
      double s2 = lj->sigma * lj->sigma / r2;
      double lj6 = s2 * s2 * s2;
      double lj12 = lj6 * lj6;
      if (e) etot += 4 * lj->epsilon * (lj12 - lj6) - shift;
      if (g) {
        double factor = -24 * lj->epsilon * (2 * lj12 - lj6) / r2;
        for (int d = 0; d < 3; d++) {
          g[3 * i + d] += dx[d] * factor;
          g[3 * j + d] -= dx[d] * factor;
        }
      }
    }
  }
  if (e) *e = etot;
}

static void ljpbc_eg_cb(const void* ctx, const double* x, double* e, double* g) {
  orc_ljpbc_energy_gradient((const orc_ljpbc*)ctx, x, e, g);
}
void orc_ljpbc_potential(const orc_ljpbc* lj, orc_pot* out) {
  out->ctx = lj;
  out->ndof = (size_t)lj->nparticle * 3;
  out->eg = ljpbc_eg_cb;
  out->constrain = NULL;
  out->convergence = 1e-6;
}

/* ------------------------------------------------------------------------- */
/* BarAndHinge                                                               */
/* ------------------------------------------------------------------------- */

static double orc_pi(void) { return 4 * atan(1); }

static int all_fixed(const orc_bah* b, const int* nodes, int nn) {
  if (!b->fixed) return 0;
  for (int k = 0; k < nn; k++)
    for (int d = 0; d < 3; d++)
      if (!b->fixed[3 * nodes[k] + d]) return 0;
  return 1;
}

/* Element order of BarAndHinge::init (src/potentials/BarAndHinge.cpp:16-67):
 * bars, then hinges, then one external-force element per node, skipping
 * elements whose DOFs are all fixed.  Kernels: stretching (:269-290, including
 * the reference's factor 2 in the gradient, :282), bending (:293-356),
 * forceEnergy (:359-369).  The substrate wall (:372-387) is not restated. */
void orc_bah_energy_gradient(const orc_bah* b, const double* x, double* e, double* g) {
  double etot = 0;
  if (g) memset(g, 0, sizeof(double) * 3 * (size_t)b->nnode);
  for (int ib = 0; ib < b->nbar; ib++) {
    const int* nd = b->bar + 2 * ib;
    if (all_fixed(b, nd, 2)) continue;
    double dx[3];
    for (int d = 0; d < 3; d++) dx[d] = x[3 * nd[0] + d] - x[3 * nd[1] + d];
    double l = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    double dl = l - b->bar_l0[ib];
    if (e) etot += 0.5 * b->bar_k[ib] * pow(dl, 2);
    if (g) {
      double gf = 2 * b->bar_k[ib] * dl / l;
      for (int d = 0; d < 3; d++) g[3 * nd[0] + d] += gf * dx[d];
      for (int d = 0; d < 3; d++) g[3 * nd[1] + d] -= gf * dx[d];
    }
  }
  const double pi = orc_pi();
  for (int ih = 0; ih < b->nhinge; ih++) {
    const int* nd = b->hinge + 4 * ih;
    if (all_fixed(b, nd, 4)) continue;
    double x1[3], x2[3], x3[3], x4[3], b1[3], b2[3], b3[3];
    for (int d = 0; d < 3; d++) {
      x1[d] = x[3 * nd[0] + d];
      x2[d] = x[3 * nd[1] + d];
      x3[d] = x[3 * nd[2] + d];
      x4[d] = x[3 * nd[3] + d];
    }
    for (int d = 0; d < 3; d++) {
      b1[d] = x2[d] - x1[d];
      b2[d] = x3[d] - x2[d];
      b3[d] = x4[d] - x3[d];
    }
    double b2m = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
    double n1[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
    double n2[3] = {b2[1] * b3[2] - b2[2] * b3[1], b2[2] * b3[0] - b2[0] * b3[2], b2[0] * b3[1] - b2[1] * b3[0]};
    double n1sq = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
    double n2sq = n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2];
    double n12m = sqrt(n1sq * n2sq);
    double c = (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) / n12m;
    c = c < -1.0 ? -1.0 : c;
    c = c > 1.0 ? 1.0 : c;
    double n1b3 = n1[0] * b3[0] + n1[1] * b3[1] + n1[2] * b3[2];
    double theta = (n1b3 >= 0) ? acos(c) : 2 * pi - acos(c);
    double dtheta = theta - b->hinge_t0[ih];
    dtheta = fmod(dtheta + pi, 2 * pi) - pi;
    if (e) etot += 0.5 * b->hinge_k[ih] * pow(dtheta, 2);
    if (g) {
      double gf = b->hinge_k[ih] * (dtheta);
      double f1 = b2m / n1sq, f2 = b2m / n2sq;
      double n1h[3] = {f1 * n1[0], f1 * n1[1], f1 * n1[2]};
      double n2h[3] = {f2 * n2[0], f2 * n2[1], f2 * n2[2]};
      double skew1 = -(b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2]) / (b2m * b2m);
      double skew2 = -(b3[0] * b2[0] + b3[1] * b2[1] + b3[2] * b2[2]) / (b2m * b2m);
      for (int d = 0; d < 3; d++) g[3 * nd[0] + d] += -gf * n1h[d];
      for (int d = 0; d < 3; d++) g[3 * nd[1] + d] += gf * ((1 - skew1) * n1h[d] - skew2 * n2h[d]);
      for (int d = 0; d < 3; d++) g[3 * nd[2] + d] += -gf * ((1 - skew2) * n2h[d] - skew1 * n1h[d]);
      for (int d = 0; d < 3; d++) g[3 * nd[3] + d] += gf * n2h[d];
    }
  }
  for (int in = 0; in < b->nnode; in++) {
    if (all_fixed(b, &in, 1)) continue;
    double fz = 0;
    if (b->force) fz = b->force_per_node ? b->force[3 * in + 2] : b->force[2];
    if (e) etot -= fz * x[3 * in + 2];
    if (g) g[3 * in + 2] -= fz;
  }
  if (e) *e = etot;
}

/* Potential::applyConstraints for single-DOF constraints (src/Potential.cpp:102-115). */
static void bah_con_cb(const void* ctx, const double* x, double* data) {
  (void)x;
  const orc_bah* b = (const orc_bah*)ctx;
  if (!b->fixed) return;
  for (int i = 0; i < 3 * b->nnode; i++)
    if (b->fixed[i]) data[i] = 0;
}
static void bah_eg_cb(const void* ctx, const double* x, double* e, double* g) {
  orc_bah_energy_gradient((const orc_bah*)ctx, x, e, g);
}
void orc_bah_potential(const orc_bah* b, orc_pot* out) {
  out->ctx = b;
  out->ndof = 3 * (size_t)b->nnode;
  out->eg = bah_eg_cb;
  out->constrain = bah_con_cb;
  out->convergence = 1e-6;
}
