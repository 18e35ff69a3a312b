/* NeighList: the connection-table container of the AGBNP3 C API.
 *
 * ABI-identical restatement of the reference header nblist.h:27-80 (same
 * struct layout, same three functions, same return codes), implemented in
 * agbnp3_amd/csrc/nblist.c and exported by libnblist.so.
 */
#ifndef LIBNBLIST_H
#define LIBNBLIST_H
#ifdef __cplusplus
extern "C" {
#endif

#define NBLIST_OK 0   /* nblist.h:27 */
#define NBLIST_ERR 1  /* nblist.h:28 */

/* nblist.h:34-36 */
typedef struct vector_ {
  double x, y, z;
} NeighVector;

/* nblist.h:39-70: rows of neighbours, neighl[i] points into neighl1 */
typedef struct neighlist_ {
  int natoms;            /* number of rows */
  int neighl_size;       /* capacity of neighl1 */
  int *nne;              /* neighbours per row */
  int *neighl1;          /* all rows, back to back */
  int **neighl;          /* row pointers into neighl1 */
  int idx_remap;         /* > 0: int2ext/ext2int are maintained */
  int *int2ext;
  int *ext2int;
  int pbc;               /* > 0: pbc_trans rows are maintained */
  NeighVector *pbc_trans1;
  NeighVector **pbc_trans;
  int data;              /* > 0: data_index rows are maintained */
  void **data_index1;
  void ***data_index;
} NeighList;

/* nblist.h:75: all fields zero / NULL (no memory is freed) */
void nblist_reset_neighbor_list(NeighList *neigh_list);
/* nblist.h:77: frees the arrays (not the struct itself) */
void nblist_delete_neighbor_list(NeighList *neigh_list);
/* nblist.h:79-80: grows rows to natoms and neighl1 to size; rebuilds the row
   pointers from nne when neighl1 moves. Returns NBLIST_OK or NBLIST_ERR. */
int nblist_reallocate_neighbor_list(NeighList *neigh_list, int natoms, int size);

#ifdef __cplusplus
}
#endif
#endif
