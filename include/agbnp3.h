/* AGBNP3 implicit-solvent energy + gradient: the drop-in C API.
 *
 * Same symbols, argument meaning, return codes and HB type constants as the
 * reference header agbnp3.h:27-112, so existing callers relink unchanged
 * (-lagbnp3 -lnblist).  Behind it, libagbnp3.so evaluates everything on one
 * MI355X with hand-written gfx950 kernels.  The README spellings
 * agbnp_initialize / agbnp_terminate / agbnp_delete (reference README.md:52,
 * 63, 146) are exported as aliases.
 */
#ifndef AGBNP3_H
#define AGBNP3_H

#ifdef __cplusplus
extern "C" {
#endif

#include "nblist.h"

#ifndef FALSE
#define FALSE 0
#endif
#ifndef TRUE
#define TRUE 1
#endif

/* agbnp3.h:43-50: interface floating type */
#ifndef FLOAT_I_H
#define FLOAT_I_H
#ifdef SINGLE_PREC_INTERFACE
typedef float float_i;
#else
typedef double float_i;
#endif
#endif

/* agbnp3.h:52-57: note OK is 2 */
#ifndef AGBNP_OK
#define AGBNP_OK (2)
#endif
#ifndef AGBNP_ERR
#define AGBNP_ERR (-1)
#endif

/* agbnp3.h:60-74: hydrogen-bond types and subtypes */
#define AGBNP_HB_INACTIVE      (0)
#define AGBNP_HB_POLARH        (1)
#define AGBNP_HB_TRIGONAL      (10)
#define AGBNP_HB_TRIGONAL_S    (30)
#define AGBNP_HB_TETRAHEDRAL   (20)
#define AGBNP_HB_TRIGONAL1     (11)
#define AGBNP_HB_TRIGONAL2     (12)
#define AGBNP_HB_TRIGONAL_OOP  (13)
#define AGBNP_HB_TETRAHEDRAL1  (21)
#define AGBNP_HB_TETRAHEDRAL2  (22)
#define AGBNP_HB_TETRAHEDRAL3  (23)

/* replaces agbnp3_initialize (agbnp3.c:67-93): idempotent; also binds the GPU */
int agbnp3_initialize(void);

/* replaces agbnp3_terminate (agbnp3.c:96-118): deletes every live instance */
void agbnp3_terminate(void);

/* replaces agbnp3_new (agbnp3.c:121-419): deep-copies all inputs (conntbl
   included), builds the Q4 spline tables and the device arena of the
   instance; returns its tag (lowest free slot) */
int agbnp3_new(int *tag, int natoms,
               float_i *x, float_i *y, float_i *z, float_i *r,
               float_i *charge, float_i dielectric_in, float_i dielectric_out,
               float_i *igamma, float_i *sgamma,
               float_i *ialpha, float_i *salpha,
               int *hbtype, float_i *hbcorr,
               int nhydrogen, int *ihydrogen,
               NeighList *conntbl, int verbose);

/* replaces agbnp3_delete (agbnp3.c:422-487); unlike the reference it frees
   each buffer exactly once */
int agbnp3_delete(int tag);

/* replaces agbnp3_ener (agbnp3.c:492-566): reads x, y, z (caller order),
   overwrites every output (caller order); `init` is ignored as in the
   reference (README.md:171) */
int agbnp3_ener(int tag, int init,
                float_i *x, float_i *y, float_i *z,
                float_i *sp, float_i *br,
                float_i *mol_volume, float_i *surf_area,
                float_i *egb, float_i (*dgbdr)[3],
                float_i *evdw, float_i *ecorr_vdw, float_i (*dvwdr)[3],
                float_i *ecav, float_i *ecorr_cav, float_i (*decav)[3],
                float_i *ehb, float_i (*dehb)[3]);

/* README spellings (reference README.md:52, 63, 146) */
int agbnp_initialize(void);
void agbnp_terminate(void);
int agbnp_delete(int tag);

#ifdef __cplusplus
}
#endif

#endif /* AGBNP3_H */
