/* libnblist: NeighList container with the semantics of the reference
 * nblist.c:30-180 (reset clears fields, delete frees arrays, reallocate
 * grows rows and storage and re-points rows from nne). */
#include <stdio.h>
#include <stdlib.h>

#include "nblist.h"

void nblist_reset_neighbor_list(NeighList *nl) {
  nl->natoms = 0;
  nl->neighl_size = 0;
  nl->nne = NULL;
  nl->neighl1 = NULL;
  nl->neighl = NULL;
  nl->idx_remap = 0;
  nl->int2ext = NULL;
  nl->ext2int = NULL;
  nl->pbc = 0;
  nl->pbc_trans1 = NULL;
  nl->pbc_trans = NULL;
  nl->data = 0;
  nl->data_index1 = NULL;
  nl->data_index = NULL;
}

void nblist_delete_neighbor_list(NeighList *nl) {
  if (!nl) return;
  free(nl->nne);
  free(nl->neighl);
  free(nl->neighl1);
  free(nl->int2ext);
  free(nl->ext2int);
  free(nl->pbc_trans);
  free(nl->pbc_trans1);
  free(nl->data_index);
  free(nl->data_index1);
}

static int grow(void **p, size_t bytes, const char *what, long count) {
  void *q = realloc(*p, bytes ? bytes : 1);
  if (!q) {
    fprintf(stderr, "nblist_reallocate_neighbor_list(): unable to (re)allocate %s (%ld entries).\n", what,
            count);
    return 0;
  }
  *p = q;
  return 1;
}

/* row pointers follow the rows' counts, as neighl1 is laid out row after row */
static void repoint(NeighList *nl) {
  long off = 0;
  for (int i = 0; i < nl->natoms; i++) {
    nl->neighl[i] = nl->neighl1 + off;
    if (nl->pbc && nl->pbc_trans) nl->pbc_trans[i] = nl->pbc_trans1 + off;
    if (nl->data && nl->data_index) nl->data_index[i] = nl->data_index1 + off;
    off += nl->nne[i];
  }
}

int nblist_reallocate_neighbor_list(NeighList *nl, int natoms, int size) {
  if (natoms > nl->natoms) {
    size_t n = (size_t)natoms;
    if (!grow((void **)&nl->nne, n * sizeof(int), "nne", natoms)) return NBLIST_ERR;
    if (!grow((void **)&nl->neighl, n * sizeof(int *), "neighl", natoms)) return NBLIST_ERR;
    for (int i = nl->natoms; i < natoms; i++) {   /* new rows start empty */
      nl->nne[i] = 0;
      nl->neighl[i] = NULL;
    }
    if (nl->idx_remap) {
      if (!grow((void **)&nl->int2ext, n * sizeof(int), "int2ext", natoms)) return NBLIST_ERR;
      if (!grow((void **)&nl->ext2int, n * sizeof(int), "ext2int", natoms)) return NBLIST_ERR;
    }
    if (nl->pbc && !grow((void **)&nl->pbc_trans, n * sizeof(NeighVector *), "pbc_trans", natoms))
      return NBLIST_ERR;
    if (nl->data && !grow((void **)&nl->data_index, n * sizeof(void **), "data_index", natoms))
      return NBLIST_ERR;
  }
  int old_rows = nl->natoms;
  if (natoms > nl->natoms) nl->natoms = natoms;
  if (size > nl->neighl_size) {
    size_t m = (size_t)size;
    if (!grow((void **)&nl->neighl1, m * sizeof(int), "neighl1", size)) return NBLIST_ERR;
    if (nl->pbc && !grow((void **)&nl->pbc_trans1, m * sizeof(NeighVector), "pbc_trans1", size))
      return NBLIST_ERR;
    if (nl->data && !grow((void **)&nl->data_index1, m * sizeof(void *), "data_index1", size))
      return NBLIST_ERR;
    nl->neighl_size = size;
    repoint(nl);
  } else if (nl->natoms > old_rows) {
    repoint(nl);
  }
  return NBLIST_OK;
}
