/* Seeded synthetic globular solutes for AGBNP3 workloads.
 *
 * A solute of `natoms` atoms is grown as a self-avoiding heavy-atom chain
 * (1.53 A bonds, 109.5 +/- 5 degree angles, uniform dihedrals) folded by
 * rejection into a sphere at 0.1 atoms/A^3, with pendant heavy atoms
 * (carbonyl / hydroxyl O, sulfone O pairs) and hydrogens at 1.0 A that bring
 * the hydrogen fraction to ~50%.  Elements, radii, charges, non-polar and
 * hydrogen-bond parameters follow the globular-solute recipe; the bond graph is
 * returned as a symmetric connection table (NeighList layout: nne + flat
 * neighbour array).  The random stream is xoshiro256** seeded by SplitMix64,
 * with Box-Muller normals, so a seed gives the same solute on every host.
 *
 * Atoms are emitted in "residue" order (each backbone atom followed by its
 * pendants and hydrogens), so callers exercise the heavy-first remapping.
 */
#ifndef AGBNP3_SYNTH_H
#define AGBNP3_SYNTH_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* info[] slots written by agbsynth_generate */
#define AGBSYNTH_INFO_SPHERE_R   0 /* final confining sphere radius (A) */
#define AGBSYNTH_INFO_MIN_NB     1 /* non-bonded heavy-heavy exclusion used (A) */
#define AGBSYNTH_INFO_BACKTRACKS 2 /* number of growth backtracks */
#define AGBSYNTH_INFO_NHEAVY     3 /* heavy atoms */
#define AGBSYNTH_INFO_NSEG       4 /* chain segments */
#define AGBSYNTH_INFO_NSIZE      5

/* Returns 0 on success, -1 on bad arguments, -2 if growth failed,
   -3 if neighl1_cap is too small. All arrays have natoms entries except
   neighl1 (neighl1_cap). ishydrogen[i] is 1 for hydrogens. */
int agbsynth_generate(uint64_t seed, int natoms,
                      double *x, double *y, double *z, double *r,
                      double *charge, double *igamma, double *sgamma,
                      double *ialpha, double *salpha,
                      int *hbtype, double *hbcorr, int *ishydrogen,
                      int *nne, int *neighl1, int neighl1_cap,
                      int *nconn_out, double *info);

/* Gaussian coordinate perturbation used by the batched / trajectory
   workloads: x += N(0, sigma) per coordinate, from its own seeded stream. */
void agbsynth_perturb(uint64_t seed, int natoms, double sigma,
                      double *x, double *y, double *z);

/* Log-uniform instance sizes in [nmin, nmax] from a seeded stream. */
void agbsynth_sizes(uint64_t seed, int ninst, int nmin, int nmax, int *sizes);

#ifdef __cplusplus
}
#endif
#endif
