/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see om_internal.h).
 *
 * Restatement of the mesh path: the msh 2.2 element-to-face step of
 * src/mesh/mesh_reading_module.F90:1122-1321, build_mesh
 * (src/mesh/mesh_connectivity_module.F90:12-59, one rank, with the periodic
 * stitch and stencil of :44-54, 976-1328 when asked for) and
 * compute_geometry_mesh (src/mesh/mesh_geometry_module.F90:16-57).
 * Loop orders and expression forms follow the reference so that ids come
 * out identical and geometry is reproduced operation for operation.
 * norm2 of a 3-vector is taken as sqrt(x*x + y*y + z*z).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

/* ------------------------------------------------------------------ arena */
typedef struct om_chunk {
  struct om_chunk *next;
  size_t used, cap;
  unsigned char *data;
} om_chunk;

void *om_arena_alloc(om_mesh *m, size_t bytes) {
  bytes = (bytes + 15u) & ~(size_t)15u;
  om_chunk *c = (om_chunk *)m->arena;
  if (!c || c->used + bytes > c->cap) {
    size_t cap = bytes > (size_t)(64u << 20) ? bytes : (size_t)(64u << 20);
    om_chunk *n = (om_chunk *)malloc(sizeof(om_chunk));
    n->data = (unsigned char *)malloc(cap);
    n->cap = cap;
    n->used = 0;
    n->next = c;
    m->arena = n;
    c = n;
  }
  void *p = c->data + c->used;
  c->used += bytes;
  return p;
}

static int *ialloc(om_mesh *m, long n) {
  int *p = (int *)om_arena_alloc(m, sizeof(int) * (size_t)(n + 1));
  memset(p, 0, sizeof(int) * (size_t)(n + 1));
  return p;
}
static double *dalloc(om_mesh *m, long n) {
  double *p = (double *)om_arena_alloc(m, sizeof(double) * (size_t)(n + 1));
  memset(p, 0, sizeof(double) * (size_t)(n + 1));
  return p;
}
static int **palloc(om_mesh *m, long n) {
  int **p = (int **)om_arena_alloc(m, sizeof(int *) * (size_t)(n + 1));
  memset(p, 0, sizeof(int *) * (size_t)(n + 1));
  return p;
}

void om_mesh_free(om_mesh *m) {
  if (!m) return;
  om_chunk *c = (om_chunk *)m->arena;
  while (c) {
    om_chunk *n = c->next;
    free(c->data);
    free(c);
    c = n;
  }
  free(m);
}

/* gmsh 2.2 element kinds used by the reader: mesh_reading_module.F90:1007-1018 */
static int n_vert_kind(int k) {
  switch (k) {
    case 4: return 4;
    case 5: return 8;
    case 6: return 6;
    case 7: return 5;
    default: return 0;
  }
}
static int n_face_kind(int k) {
  switch (k) {
    case 4: return 4;
    case 5: return 6;
    case 6: return 5;
    case 7: return 5;
    default: return 0;
  }
}

/* Local face -> local vertex tables, mesh_reading_module.F90:1163-1316. */
static const int TET_F[4][4] = {{1, 3, 2, 0}, {1, 2, 4, 0}, {1, 4, 3, 0}, {2, 3, 4, 0}};
static const int TET_N[4] = {3, 3, 3, 3};
static const int HEX_F[6][4] = {{1, 4, 3, 2}, {3, 4, 8, 7}, {7, 8, 5, 6},
                                {1, 2, 6, 5}, {1, 5, 8, 4}, {2, 3, 7, 6}};
static const int HEX_N[6] = {4, 4, 4, 4, 4, 4};
static const int PRI_F[5][4] = {{1, 3, 2, 0}, {4, 5, 6, 0}, {1, 2, 5, 4}, {2, 3, 6, 5}, {1, 4, 6, 3}};
static const int PRI_N[5] = {3, 3, 4, 4, 4};
static const int PYR_F[5][4] = {{1, 4, 3, 2}, {1, 2, 5, 0}, {2, 3, 5, 0}, {3, 4, 5, 0}, {1, 5, 4, 0}};
static const int PYR_N[5] = {4, 3, 3, 3, 3};

static void local_face(int kind, int j, int *nfv, const int **lv) {
  switch (kind) {
    case 4: *nfv = TET_N[j - 1]; *lv = TET_F[j - 1]; break;
    case 5: *nfv = HEX_N[j - 1]; *lv = HEX_F[j - 1]; break;
    case 6: *nfv = PRI_N[j - 1]; *lv = PRI_F[j - 1]; break;
    default: *nfv = PYR_N[j - 1]; *lv = PYR_F[j - 1]; break;
  }
}

#define VC(m, i, d) ((m)->vcoord[3 * (i) + (d)])
#define EC(m, i, d) ((m)->ecoord[3 * (i) + (d)])
#define FC(m, i, d) ((m)->fcoord[3 * (i) + (d)])
#define FN(m, i, d) ((m)->fnorm[3 * (i) + (d)])
#define SN(m, i, d) ((m)->sf_norm[3 * (i) + (d)])

/* ------------------------------------------------- reader (element faces) */
static void read_elements(om_mesh *m, int nv, const double *coord, int ne, const int *kind,
                          const int *evert, int nbf, const int *bf_nv, const int *bf_vert,
                          const int *bf_tag) {
  m->nv = nv;
  m->vcoord = dalloc(m, 3L * (nv + 1));
  for (int i = 1; i <= nv; i++)
    for (int d = 0; d < 3; d++) VC(m, i, d) = coord[3 * (i - 1) + d];

  m->ne = ne;
  m->e_kind = ialloc(m, ne);
  m->e_nv = ialloc(m, ne);
  m->e_nf = ialloc(m, ne);
  m->e_vert = palloc(m, ne);
  m->e_face = palloc(m, ne);
  long nmf = 0;
  for (int i = 1; i <= ne; i++) nmf += n_face_kind(kind[i - 1]);
  m->nmf = (int)nmf;
  m->mf_nv = ialloc(m, nmf);
  m->mf_left = ialloc(m, nmf);
  m->mf_leftf = ialloc(m, nmf);
  m->mf_right = ialloc(m, nmf);
  m->mf_rightf = ialloc(m, nmf);
  m->mf_vert = palloc(m, nmf);

  long pos = 0;
  int n_faces = 1;
  for (int i = 1; i <= ne; i++) {
    int k = kind[i - 1];
    int nvk = n_vert_kind(k), nfk = n_face_kind(k);
    m->e_kind[i] = k;
    m->e_nv[i] = nvk;
    m->e_nf[i] = nfk;
    m->e_vert[i] = ialloc(m, nvk);
    for (int j = 1; j <= nvk; j++) m->e_vert[i][j] = evert[pos + j - 1];
    pos += nvk;
    m->e_face[i] = ialloc(m, nfk);
    for (int j = 1; j <= nfk; j++) {
      int nfv;
      const int *lv;
      local_face(k, j, &nfv, &lv);
      int f = n_faces + j - 1;
      m->mf_left[f] = i;
      m->mf_leftf[f] = j;
      m->mf_nv[f] = nfv;
      m->mf_vert[f] = ialloc(m, nfv);
      for (int q = 0; q < nfv; q++) m->mf_vert[f][q + 1] = m->e_vert[i][lv[q]];
      m->e_face[i][j] = f;
    }
    n_faces += nfk;
  }

  m->nbf = nbf;
  m->bf_nv = ialloc(m, nbf);
  m->bf_tag = ialloc(m, nbf);
  m->bf_vert = palloc(m, nbf);
  pos = 0;
  for (int i = 1; i <= nbf; i++) {
    m->bf_nv[i] = bf_nv[i - 1];
    m->bf_tag[i] = bf_tag[i - 1];
    m->bf_vert[i] = ialloc(m, bf_nv[i - 1]);
    for (int j = 1; j <= bf_nv[i - 1]; j++) m->bf_vert[i][j] = bf_vert[pos + j - 1];
    pos += bf_nv[i - 1];
  }
}

/* build_node_elem_neigh, mesh_connectivity_module.F90:475-509 */
static void build_node_elem_neigh(om_mesh *m) {
  m->v_nen = ialloc(m, m->nv);
  m->v_elem = palloc(m, m->nv);
  for (int i = 1; i <= m->ne; i++)
    for (int j = 1; j <= m->e_nv[i]; j++) m->v_nen[m->e_vert[i][j]]++;
  for (int i = 1; i <= m->nv; i++) {
    m->v_elem[i] = ialloc(m, m->v_nen[i]);
    m->v_nen[i] = 0;
  }
  for (int i = 1; i <= m->ne; i++)
    for (int j = 1; j <= m->e_nv[i]; j++) {
      int v = m->e_vert[i][j];
      m->v_elem[v][++m->v_nen[v]] = i;
    }
}

/* renumber_vert_mesh, mesh_connectivity_module.F90:851-926 (breadth-first
 * over vertex -> element -> vertex, restarting at the lowest unvisited id) */
static void renumber_vert_mesh(om_mesh *m) {
  int nv = m->nv;
  int *otn = (int *)malloc(sizeof(int) * (nv + 1));
  int *nto = ialloc(m, nv);
  for (int i = 1; i <= nv; i++) otn[i] = nto[i] = -1;
  otn[1] = 1;
  nto[1] = 1;
  int iloc = 2;
  int iloc_start = iloc - 1, iloc_end = iloc - 1;
  while (iloc <= nv) {
    int iloc_before = iloc;
    for (int i = iloc_start; i <= iloc_end; i++) {
      int v = nto[i];
      for (int j = 1; j <= m->v_nen[v]; j++) {
        int e = m->v_elem[v][j];
        for (int k = 1; k <= m->e_nv[e]; k++) {
          int w = m->e_vert[e][k];
          if (otn[w] < 0) {
            otn[w] = iloc;
            nto[iloc] = w;
            iloc++;
          }
        }
      }
    }
    if (iloc == iloc_before) {
      for (int i = 1; i <= nv; i++)
        if (otn[i] < 0) {
          otn[i] = iloc;
          nto[iloc] = i;
          iloc++;
          break;
        }
    }
    iloc_start = iloc_end + 1;
    iloc_end = iloc - 1;
  }
  /* vert(otn(i)) = vert_tmp(i): move coordinates and element lists */
  double *nc = dalloc(m, 3L * (nv + 1));
  int *nnen = ialloc(m, nv);
  int **nel = palloc(m, nv);
  for (int i = 1; i <= nv; i++) {
    int n = otn[i];
    for (int d = 0; d < 3; d++) nc[3 * n + d] = VC(m, i, d);
    nnen[n] = m->v_nen[i];
    nel[n] = m->v_elem[i];
  }
  m->vcoord = nc;
  m->v_nen = nnen;
  m->v_elem = nel;
  for (int i = 1; i <= m->ne; i++)
    for (int j = 1; j <= m->e_nv[i]; j++) m->e_vert[i][j] = otn[m->e_vert[i][j]];
  for (int i = 1; i <= m->nmf; i++)
    for (int j = 1; j <= m->mf_nv[i]; j++) m->mf_vert[i][j] = otn[m->mf_vert[i][j]];
  for (int i = 1; i <= m->nbf; i++)
    for (int j = 1; j <= m->bf_nv[i]; j++) m->bf_vert[i][j] = otn[m->bf_vert[i][j]];
  m->vert_nto = nto;
  free(otn);
}

/* renumber_elem_mesh, mesh_connectivity_module.F90:752-849 (one rank):
 * first appearance over the vertices' element lists */
static void renumber_elem_mesh(om_mesh *m) {
  int ne = m->ne;
  int *otn = (int *)malloc(sizeof(int) * (ne + 1));
  for (int i = 1; i <= ne; i++) otn[i] = -1;
  int iloc = 1;
  for (int i = 1; i <= m->nv; i++)
    for (int j = 1; j <= m->v_nen[i]; j++) {
      int le = m->v_elem[i][j];
      if (otn[le] <= 0) otn[le] = iloc++;
    }
  int *kind = ialloc(m, ne), *nvv = ialloc(m, ne), *nf = ialloc(m, ne);
  int **vert = palloc(m, ne), **face = palloc(m, ne);
  int *nto = ialloc(m, ne);
  for (int i = 1; i <= ne; i++) {
    int n = otn[i];
    kind[n] = m->e_kind[i];
    nvv[n] = m->e_nv[i];
    nf[n] = m->e_nf[i];
    vert[n] = m->e_vert[i];
    face[n] = m->e_face[i];
    nto[n] = i;
  }
  m->e_kind = kind;
  m->e_nv = nvv;
  m->e_nf = nf;
  m->e_vert = vert;
  m->e_face = face;
  m->elem_nto = nto;
  for (int i = 1; i <= m->nmf; i++) m->mf_left[i] = otn[m->mf_left[i]];
  for (int i = 1; i <= m->nv; i++)
    for (int j = 1; j <= m->v_nen[i]; j++) m->v_elem[i][j] = otn[m->v_elem[i][j]];
  free(otn);
}

/* same_faces / same_faces_boundary, mesh_connectivity_module.F90:679-724 */
static int same_vert_sets(int na, const int *a, int nb, const int *b) {
  if (na != nb) return 0;
  int same = 1;
  for (int i = 1; i <= na; i++) {
    int found = 0;
    for (int j = 1; j <= nb; j++)
      if (a[i] == b[j]) found = 1;
    same = found && same;
  }
  return same;
}

/* build_faces, mesh_connectivity_module.F90:61-211: de-duplicate the element
 * faces in buckets keyed by their smallest vertex id, insertion order kept */
static void build_faces(om_mesh *m) {
  int nmf = m->nmf;
  int id_glob_max = 0;
  int *fmin = (int *)malloc(sizeof(int) * (nmf + 1));
  for (int j = 1; j <= nmf; j++) {
    int mn = m->mf_vert[j][1];
    for (int k = 2; k <= m->mf_nv[j]; k++)
      if (m->mf_vert[j][k] < mn) mn = m->mf_vert[j][k];
    fmin[j] = mn;
    if (mn > id_glob_max) id_glob_max = mn;
  }
  int *ncol = (int *)calloc((size_t)id_glob_max + 1, sizeof(int));
  for (int j = 1; j <= nmf; j++) ncol[fmin[j]]++;
  /* bucket j holds slots [bstart[j], bstart[j] + ncol[j]) of a flat table */
  long *bstart = (long *)malloc(sizeof(long) * ((size_t)id_glob_max + 2));
  bstart[1] = 0;
  for (int j = 1; j <= id_glob_max; j++) bstart[j + 1] = bstart[j] + ncol[j];
  long nslot = bstart[id_glob_max + 1];
  int *s_nv = (int *)malloc(sizeof(int) * (size_t)(nslot + 1));
  int *s_src = (int *)malloc(sizeof(int) * (size_t)(nslot + 1)); /* minimal face giving verts */
  int *s_left = (int *)malloc(sizeof(int) * (size_t)(nslot + 1));
  int *s_leftf = (int *)malloc(sizeof(int) * (size_t)(nslot + 1));
  int *s_right = (int *)calloc((size_t)nslot + 1, sizeof(int));
  int *s_rightf = (int *)calloc((size_t)nslot + 1, sizeof(int));
  for (long s = 0; s < nslot; s++) s_nv[s] = -1;

  int n_faces = 0;
  for (int j = 1; j <= nmf; j++) {
    int b = fmin[j];
    int contained = 0;
    for (int k = 0; k < ncol[b]; k++) {
      long s = bstart[b] + k;
      if (s_nv[s] > 0) {
        int src = s_src[s];
        if (same_vert_sets(m->mf_nv[j], m->mf_vert[j], m->mf_nv[src], m->mf_vert[src])) {
          contained = 1;
          s_right[s] = m->mf_left[j];
          s_rightf[s] = m->mf_leftf[j];
          break;
        }
      }
    }
    if (!contained) {
      for (int k = 0; k < ncol[b]; k++) {
        long s = bstart[b] + k;
        if (s_nv[s] <= 0) {
          n_faces++;
          s_nv[s] = m->mf_nv[j];
          s_src[s] = j;
          s_left[s] = m->mf_left[j];
          s_leftf[s] = m->mf_leftf[j];
          break;
        }
      }
    }
  }
  /* boundary faces: the last matching listing sets the tag */
  for (int j = 1; j <= m->nbf; j++) {
    int mn = m->bf_vert[j][1];
    for (int k = 2; k <= m->bf_nv[j]; k++)
      if (m->bf_vert[j][k] < mn) mn = m->bf_vert[j][k];
    if (mn > id_glob_max) continue;
    for (int k = 0; k < ncol[mn]; k++) {
      long s = bstart[mn] + k;
      if (s_nv[s] > 0) {
        int src = s_src[s];
        if (same_vert_sets(m->bf_nv[j], m->bf_vert[j], m->mf_nv[src], m->mf_vert[src])) {
          if (s_right[s] <= 0) {
            s_right[s] = -m->bf_tag[j];
            s_rightf[s] = 0;
          }
          break;
        }
      }
    }
  }

  m->nf = n_faces;
  m->f_nv = ialloc(m, n_faces);
  m->f_left = ialloc(m, n_faces);
  m->f_leftf = ialloc(m, n_faces);
  m->f_right = ialloc(m, n_faces);
  m->f_rightf = ialloc(m, n_faces);
  m->f_vert = palloc(m, n_faces);
  m->farea = dalloc(m, n_faces);
  m->fnorm = dalloc(m, 3L * (n_faces + 1));
  m->fcoord = dalloc(m, 3L * (n_faces + 1));
  n_faces = 0;
  for (int b = 1; b <= id_glob_max; b++)
    for (int k = 0; k < ncol[b]; k++) {
      long s = bstart[b] + k;
      if (s_nv[s] > 0) {
        n_faces++;
        int src = s_src[s];
        m->f_nv[n_faces] = s_nv[s];
        m->f_vert[n_faces] = ialloc(m, s_nv[s]);
        for (int i = 1; i <= s_nv[s]; i++) m->f_vert[n_faces][i] = m->mf_vert[src][i];
        m->f_left[n_faces] = s_left[s];
        m->f_leftf[n_faces] = s_leftf[s];
        m->f_right[n_faces] = s_right[s];
        m->f_rightf[n_faces] = s_rightf[s];
        m->e_face[s_left[s]][s_leftf[s]] = n_faces;
        if (s_right[s] > 0) m->e_face[s_right[s]][s_rightf[s]] = n_faces;
      }
    }
  free(fmin);
  free(ncol);
  free(bstart);
  free(s_nv);
  free(s_src);
  free(s_left);
  free(s_leftf);
  free(s_right);
  free(s_rightf);
}

/* build_node_face_connectivity, mesh_connectivity_module.F90:299-333 */
static void build_node_face_connectivity(om_mesh *m) {
  m->v_nfn = ialloc(m, m->nv);
  m->v_face = palloc(m, m->nv);
  for (int i = 1; i <= m->nf; i++)
    for (int j = 1; j <= m->f_nv[i]; j++) m->v_nfn[m->f_vert[i][j]]++;
  int *fill = ialloc(m, m->nv);
  for (int i = 1; i <= m->nv; i++) m->v_face[i] = ialloc(m, m->v_nfn[i]);
  for (int i = 1; i <= m->nf; i++)
    for (int j = 1; j <= m->f_nv[i]; j++) {
      int v = m->f_vert[i][j];
      m->v_face[v][++fill[v]] = i;
    }
}

/* renumber_faces_mesh, mesh_connectivity_module.F90:928-974 */
static void renumber_faces_mesh(om_mesh *m) {
  int nf = m->nf;
  int *otn = (int *)malloc(sizeof(int) * (nf + 1));
  for (int i = 1; i <= nf; i++) otn[i] = -1;
  int iloc = 1;
  for (int i = 1; i <= m->nv; i++)
    for (int j = 1; j <= m->v_nfn[i]; j++) {
      int f = m->v_face[i][j];
      if (otn[f] < 0) otn[f] = iloc++;
    }
  int *nv_ = ialloc(m, nf), *l = ialloc(m, nf), *lf = ialloc(m, nf), *r = ialloc(m, nf),
      *rf = ialloc(m, nf);
  int **vv = palloc(m, nf);
  for (int i = 1; i <= nf; i++) {
    int n = otn[i];
    nv_[n] = m->f_nv[i];
    l[n] = m->f_left[i];
    lf[n] = m->f_leftf[i];
    r[n] = m->f_right[i];
    rf[n] = m->f_rightf[i];
    vv[n] = m->f_vert[i];
  }
  m->f_nv = nv_;
  m->f_left = l;
  m->f_leftf = lf;
  m->f_right = r;
  m->f_rightf = rf;
  m->f_vert = vv;
  for (int i = 1; i <= m->ne; i++)
    for (int j = 1; j <= m->e_nf[i]; j++) m->e_face[i][j] = otn[m->e_face[i][j]];
  for (int i = 1; i <= m->nv; i++)
    for (int j = 1; j <= m->v_nfn[i]; j++) m->v_face[i][j] = otn[m->v_face[i][j]];
  free(otn);
}

/* build_sub_faces, mesh_connectivity_module.F90:335-402 (vertex-major ids) */
static void build_sub_faces(om_mesh *m) {
  long nsf = 0;
  m->f_nsf = ialloc(m, m->nf);
  m->f_sf = palloc(m, m->nf);
  for (int i = 1; i <= m->nf; i++) {
    m->f_nsf[i] = m->f_nv[i];
    m->f_sf[i] = ialloc(m, m->f_nv[i]);
    nsf += m->f_nv[i];
  }
  m->nsf = (int)nsf;
  m->sf_vert = ialloc(m, nsf);
  m->sf_face = ialloc(m, nsf);
  m->sf_loc = ialloc(m, nsf);
  m->sf_le = ialloc(m, nsf);
  m->sf_re = ialloc(m, nsf);
  m->sf_lse = ialloc(m, nsf);
  m->sf_rse = ialloc(m, nsf);
  m->sf_area = dalloc(m, nsf);
  m->sf_norm = dalloc(m, 3 * (nsf + 1));
  m->e_nsf = ialloc(m, m->ne);
  m->e_sf = palloc(m, m->ne);
  for (int i = 1; i <= m->ne; i++) {
    int n = 0;
    for (int j = 1; j <= m->e_nf[i]; j++) n += m->f_nv[m->e_face[i][j]];
    m->e_nsf[i] = n;
    m->e_sf[i] = ialloc(m, n);
  }
  int *ffill = ialloc(m, m->nf), *efill = ialloc(m, m->ne);
  int id = 1;
  for (int i = 1; i <= m->nv; i++)
    for (int j = 1; j <= m->v_nfn[i]; j++) {
      int f = m->v_face[i][j];
      m->f_sf[f][++ffill[f]] = id;
      m->sf_face[id] = f;
      m->sf_vert[id] = i;
      m->sf_le[id] = m->f_left[f];
      m->sf_re[id] = m->f_right[f];
      int l = m->f_left[f];
      m->e_sf[l][++efill[l]] = id;
      int r = m->f_right[f];
      if (r > 0) m->e_sf[r][++efill[r]] = id;
      id++;
    }
}

/* build_sub_elems, mesh_connectivity_module.F90:404-452 (cell-major ids) */
static void build_sub_elems(om_mesh *m) {
  long nse = 0;
  m->e_nse = ialloc(m, m->ne);
  m->e_se = palloc(m, m->ne);
  for (int i = 1; i <= m->ne; i++) {
    m->e_nse[i] = m->e_nv[i];
    m->e_se[i] = ialloc(m, m->e_nv[i]);
    nse += m->e_nv[i];
  }
  m->nse = (int)nse;
  m->se_loc = ialloc(m, nse);
  m->se_vert = ialloc(m, nse);
  m->se_elem = ialloc(m, nse);
  m->se_nsf = ialloc(m, nse);
  m->se_sf = palloc(m, nse);
  m->se_vol = dalloc(m, nse);
  int id = 1;
  for (int i = 1; i <= m->ne; i++)
    for (int j = 1; j <= m->e_nv[i]; j++) {
      int n = 0;
      for (int k = 1; k <= m->e_nsf[i]; k++)
        if (m->sf_vert[m->e_sf[i][k]] == m->e_vert[i][j]) n++;
      m->se_nsf[id] = n;
      m->se_sf[id] = ialloc(m, n);
      id++;
    }
  id = 1;
  for (int i = 1; i <= m->ne; i++)
    for (int j = 1; j <= m->e_nv[i]; j++) {
      int loc = 1;
      m->e_se[i][j] = id;
      for (int k = 1; k <= m->e_nsf[i]; k++) {
        int sf = m->e_sf[i][k];
        if (m->sf_vert[sf] == m->e_vert[i][j]) {
          m->se_elem[id] = i;
          m->se_vert[id] = m->sf_vert[sf];
          m->se_sf[id][loc++] = sf;
        }
      }
      id++;
    }
}

/* find_sub_elems_around_sub_face, mesh_connectivity_module.F90:454-473 */
static void find_sub_elems_around_sub_face(om_mesh *m) {
  for (int i = 1; i <= m->nse; i++) {
    int e = m->se_elem[i];
    for (int j = 1; j <= m->se_nsf[i]; j++) {
      int sf = m->se_sf[i][j];
      if (m->sf_le[sf] == e) m->sf_lse[sf] = i;
      if (m->sf_re[sf] == e) m->sf_rse[sf] = i;
    }
  }
}

/* build_node_sub_face_connectivity, mesh_connectivity_module.F90:511-556 */
static void build_node_sub_face_connectivity(om_mesh *m) {
  m->v_nsfn = ialloc(m, m->nv);
  m->v_nsen = ialloc(m, m->nv);
  m->v_sf = palloc(m, m->nv);
  m->v_se = palloc(m, m->nv);
  for (int i = 1; i <= m->nsf; i++) m->v_nsfn[m->sf_vert[i]]++;
  for (int i = 1; i <= m->nse; i++) m->v_nsen[m->se_vert[i]]++;
  for (int i = 1; i <= m->nv; i++) {
    m->v_sf[i] = ialloc(m, m->v_nsfn[i]);
    m->v_se[i] = ialloc(m, m->v_nsen[i]);
    m->v_nsfn[i] = 0;
    m->v_nsen[i] = 0;
  }
  for (int i = 1; i <= m->nsf; i++) {
    int v = m->sf_vert[i];
    m->sf_loc[i] = m->v_nsfn[v] + 1;
    m->v_sf[v][++m->v_nsfn[v]] = i;
  }
  for (int i = 1; i <= m->nse; i++) {
    int v = m->se_vert[i];
    m->se_loc[i] = m->v_nsen[v] + 1;
    m->v_se[v][++m->v_nsen[v]] = i;
  }
}

/* compute_elem_tag, mesh_connectivity_module.F90:558-579 */
static void compute_elem_tag(om_mesh *m) {
  m->e_tag = ialloc(m, m->ne);
  for (int i = 1; i <= m->ne; i++) m->e_tag[i] = -1;
  for (int i = 1; i <= m->nf; i++)
    if (m->f_right[i] < 0) {
      int tag = -m->f_right[i];
      int e = m->f_left[i];
      if (m->e_tag[e] < 0)
        m->e_tag[e] = tag;
      else if (m->e_tag[e] > tag)
        m->e_tag[e] = tag;
    }
}

/* compute_neigh_by_vert, mesh_connectivity_module.F90:595-654 */
static void compute_neigh_by_vert(om_mesh *m) {
  m->e_nnbv = ialloc(m, m->ne);
  m->e_nbv = palloc(m, m->ne);
  int cap = 64;
  int *tmp = (int *)malloc(sizeof(int) * (cap + 1));
  for (int i = 1; i <= m->ne; i++) {
    int n = 0;
    for (int j = 1; j <= m->e_nv[i]; j++) {
      int v = m->e_vert[i][j];
      for (int k = 1; k <= m->v_nen[v]; k++)
        if (m->v_elem[v][k] != i) n++;
    }
    if (n > cap) {
      cap = 2 * n;
      tmp = (int *)realloc(tmp, sizeof(int) * (cap + 1));
    }
    int iloc = 1;
    for (int j = 1; j <= m->e_nv[i]; j++) {
      int v = m->e_vert[i][j];
      for (int k = 1; k <= m->v_nen[v]; k++)
        if (m->v_elem[v][k] != i) tmp[iloc++] = m->v_elem[v][k];
    }
    int nn = n;
    for (int j = 1; j <= n; j++)
      for (int k = j + 1; k <= n; k++)
        if (tmp[k] > 0 && tmp[j] == tmp[k]) {
          tmp[k] = -1;
          nn--;
        }
    m->e_nnbv[i] = nn;
    m->e_nbv[i] = ialloc(m, nn);
    iloc = 1;
    for (int j = 1; j <= n; j++)
      if (tmp[j] > 0) m->e_nbv[i][iloc++] = tmp[j];
  }
  free(tmp);
}

/* find_if_vert_is_bound, mesh_geometry_module.F90:714-738 (the same test is
 * run by build_mesh before the face normals exist; the geometry pass reruns
 * it, so only the final result is kept here) */
static void find_if_vert_is_bound(om_mesh *m, int b2d) {
  if (!m->vbound) m->vbound = ialloc(m, m->nv);
  for (int i = 1; i <= m->nv; i++) {
    m->vbound[i] = 0;
    for (int j = 1; j <= m->v_nfn[i]; j++) {
      int f = m->v_face[i][j];
      if (b2d) {
        if (m->f_right[f] <= 0 && fabs(FN(m, f, 2)) < 1e-12) m->vbound[i] = 1;
      } else {
        if (m->f_right[f] <= 0) m->vbound[i] = 1;
      }
    }
  }
}

/* ----------------------------------------------------------- periodic */
#define OM_PTOL 1e-12
static int on_plane(double x, double c) { return x < c + OM_PTOL && x > c - OM_PTOL; }

/* stitch_periodic_faces, mesh_connectivity_module.F90:976-1073, with the
 * face centres taken as vertex averages as build_mesh does first (:44-52):
 * faces on x = -0.5 and x = 0.5 (then y, then z) whose other two centre
 * coordinates agree within 1e-12 become each other's right neighbours. */
static int stitch_periodic_faces(om_mesh *m) {
  double *fc = (double *)malloc(sizeof(double) * 3 * (size_t)(m->nf + 1));
  for (int i = 1; i <= m->nf; i++) {
    for (int d = 0; d < 3; d++) fc[3 * i + d] = 0.0;
    for (int j = 1; j <= m->f_nv[i]; j++)
      for (int d = 0; d < 3; d++) fc[3 * i + d] = fc[3 * i + d] + VC(m, m->f_vert[i][j], d);
    for (int d = 0; d < 3; d++) fc[3 * i + d] = fc[3 * i + d] / (double)m->f_nv[i];
  }
  int cnt[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int i = 1; i <= m->nf; i++)
    for (int d = 0; d < 3; d++) {
      if (on_plane(fc[3 * i + d], -0.5)) cnt[d][0]++;
      if (on_plane(fc[3 * i + d], 0.5)) cnt[d][1]++;
    }
  for (int d = 0; d < 3; d++)
    if (cnt[d][0] != cnt[d][1]) { /* "Error periodic boundaries !" */
      free(fc);
      return 0;
    }
  int *lst[3][2];
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 2; k++) {
      lst[d][k] = (int *)malloc(sizeof(int) * (size_t)(cnt[d][k] + 1));
      cnt[d][k] = 0;
    }
  for (int i = 1; i <= m->nf; i++)
    for (int d = 0; d < 3; d++) {
      if (on_plane(fc[3 * i + d], -0.5)) lst[d][0][++cnt[d][0]] = i;
      if (on_plane(fc[3 * i + d], 0.5)) lst[d][1][++cnt[d][1]] = i;
    }
  for (int d = 0; d < 3; d++) {
    const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;  /* the two other axes */
    for (int i = 1; i <= cnt[d][0]; i++)
      for (int j = 1; j <= cnt[d][1]; j++) {
        int ii = lst[d][0][i], jj = lst[d][1][j];
        double da = fc[3 * ii + a] - fc[3 * jj + a], db = fc[3 * ii + b] - fc[3 * jj + b];
        if (sqrt(da * da + db * db) < OM_PTOL) {
          m->f_right[ii] = m->f_left[jj];
          m->f_right[jj] = m->f_left[ii];
        }
      }
  }
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 2; k++) free(lst[d][k]);
  free(fc);
  for (int i = 1; i <= m->nsf; i++) {
    m->sf_le[i] = m->f_left[m->sf_face[i]];
    m->sf_re[i] = m->f_right[m->sf_face[i]];
  }
  return 1;
}

/* two sub-faces join the same pair of cells, in either orientation */
static int same_cells(const om_mesh *m, int a, int b) {
  return (m->sf_le[a] == m->sf_re[b] && m->sf_re[a] == m->sf_le[b]) ||
         (m->sf_le[a] == m->sf_le[b] && m->sf_re[a] == m->sf_re[b]);
}

/* process_pair of build_periodic_stencil, mesh_connectivity_module.F90:1223-1326:
 * v1 takes v2 as its periodic partner.  Returns 0 when the extended list
 * cannot be filled (a case the reference leaves undefined). */
static int process_pair(om_mesh *m, int v1, int v2) {
  /* step 1: right sub-element of v1's stitched sub-faces, from v2's list */
  for (int ki = 1; ki <= m->v_nsfn[v1]; ki++) {
    int sfki = m->v_sf[v1][ki], re = m->sf_re[sfki];
    if (re > 0 && m->sf_rse[sfki] == 0)
      for (int k = 1; k <= m->v_nsen[v2]; k++) {
        int id_se = m->v_se[v2][k];
        if (m->se_elem[id_se] == re) {
          m->sf_rse[sfki] = id_se;
          break;
        }
      }
  }
  /* step 2: v2's sub-faces appended to v1's list, without repeats */
  int n1 = m->v_nsfn[v1], n2 = m->v_nsfn[v2], nm = 0;
  for (int ki = 1; ki <= n1; ki++)
    for (int kii = 1; kii <= n2; kii++)
      if (same_cells(m, m->v_sf[v1][ki], m->v_sf[v2][kii])) nm++;
  int st = n1 + n2 - nm;
  if (st < n1) return 0;
  int *nl = ialloc(m, st);
  for (int ki = 1; ki <= n1; ki++) nl[ki] = m->v_sf[v1][ki];
  for (int ki = n1 + 1; ki <= st; ki++) nl[ki] = -1;
  for (int ki = 1; ki <= st; ki++) {
    if (nl[ki] >= 0) continue;
    for (int kii = 1; kii <= n2; kii++) {
      int sfkii = m->v_sf[v2][kii], found = 0;
      for (int p = 1; p <= ki - 1; p++) {
        if (nl[p] < 0) continue;
        if (same_cells(m, nl[p], sfkii)) found = 1;
      }
      if (!found) {
        nl[ki] = sfkii;
        break;
      }
    }
    if (nl[ki] < 0) return 0;
  }
  m->v_sf[v1] = nl;
  m->v_nsfn[v1] = st;
  /* step 3: v2's own sub-elements appended to v1's list */
  int ne1 = m->v_nsen[v1], add = 0;
  for (int k = 1; k <= m->v_nsen[v2]; k++)
    if (m->se_vert[m->v_se[v2][k]] == v2) add++;
  int *el = ialloc(m, ne1 + add);
  for (int k = 1; k <= ne1; k++) el[k] = m->v_se[v1][k];
  add = 0;
  for (int k = 1; k <= m->v_nsen[v2]; k++) {
    int id_se = m->v_se[v2][k];
    if (m->se_vert[id_se] == v2) el[ne1 + ++add] = id_se;
  }
  m->v_se[v1] = el;
  m->v_nsen[v1] = ne1 + add;
  return 1;
}

/* build_periodic_stencil, mesh_connectivity_module.F90:1075-1221: vertex
 * pairs across x, y and z, then the xy, xz and yz edge pairs and the xyz
 * corner pair, each processed both ways, in that order */
static int build_periodic_stencil(om_mesh *m) {
  int *lst[3][2], cnt[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 2; k++) lst[d][k] = (int *)malloc(sizeof(int) * (size_t)(m->nv + 1));
  for (int i = 1; i <= m->nv; i++)
    for (int d = 0; d < 3; d++) {
      if (on_plane(VC(m, i, d), -0.5)) lst[d][0][++cnt[d][0]] = i;
      if (on_plane(VC(m, i, d), 0.5)) lst[d][1][++cnt[d][1]] = i;
    }
  int ok = 1;
#define PAIR(a, b)                                              \
  do {                                                          \
    ok = ok && process_pair(m, (a), (b)) && process_pair(m, (b), (a)); \
  } while (0)
  for (int d = 0; d < 3; d++) {
    const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
    for (int i = 1; i <= cnt[d][0]; i++)
      for (int j = 1; j <= cnt[d][1]; j++) {
        int ii = lst[d][0][i], jj = lst[d][1][j];
        double da = VC(m, ii, a) - VC(m, jj, a), db = VC(m, ii, b) - VC(m, jj, b);
        if (sqrt(da * da + db * db) < OM_PTOL) PAIR(ii, jj);
      }
  }
  /* edges: (d1, d2) both on their lower planes <-> both on their upper
   * planes, matched by the remaining axis; listed from the d1 sets */
  static const int edge[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
  for (int e = 0; e < 3; e++) {
    const int d1 = edge[e][0], d2 = edge[e][1], r = edge[e][2];
    for (int i = 1; i <= cnt[d1][0]; i++) {
      int ii = lst[d1][0][i];
      if (!on_plane(VC(m, ii, d2), -0.5)) continue;
      for (int j = 1; j <= cnt[d1][1]; j++) {
        int jj = lst[d1][1][j];
        if (!on_plane(VC(m, jj, d2), 0.5)) continue;
        if (fabs(VC(m, ii, r) - VC(m, jj, r)) < OM_PTOL) PAIR(ii, jj);
      }
    }
  }
  for (int i = 1; i <= cnt[0][0]; i++) {
    int ii = lst[0][0][i];
    if (!(on_plane(VC(m, ii, 1), -0.5) && on_plane(VC(m, ii, 2), -0.5))) continue;
    for (int j = 1; j <= cnt[0][1]; j++) {
      int jj = lst[0][1][j];
      if (on_plane(VC(m, jj, 1), 0.5) && on_plane(VC(m, jj, 2), 0.5)) PAIR(ii, jj);
    }
  }
#undef PAIR
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 2; k++) free(lst[d][k]);
  return ok;
}

/* ----------------------------------------------------------- geometry */
static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = -(a[0] * b[2] - a[2] * b[0]);
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double nrm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* compute_face_cgeom / compute_face_centroid, mesh_geometry_module.F90:193-235 */
static void face_centroid(om_mesh *m, int f, double *c) {
  double cg[3] = {0, 0, 0};
  int n = m->f_nv[f];
  for (int j = 1; j <= n; j++)
    for (int d = 0; d < 3; d++) cg[d] = cg[d] + VC(m, m->f_vert[f][j], d);
  for (int d = 0; d < 3; d++) cg[d] = cg[d] / (double)n;
  double sum_area = 0.0;
  c[0] = c[1] = c[2] = 0.0;
  for (int j = 1; j <= n; j++) {
    int jp = 1 + (j % n);
    const double *cv = &VC(m, m->f_vert[f][j], 0);
    const double *cvp = &VC(m, m->f_vert[f][jp], 0);
    double a[3], b[3], x[3];
    for (int d = 0; d < 3; d++) {
      a[d] = cv[d] - cg[d];
      b[d] = cvp[d] - cg[d];
    }
    cross3(a, b, x);
    double area = 0.5 * nrm3(x);
    for (int d = 0; d < 3; d++) c[d] = c[d] + area * (cv[d] + cvp[d] + cg[d]) / 3.0;
    sum_area = sum_area + area;
  }
  for (int d = 0; d < 3; d++) c[d] = c[d] / sum_area;
}

/* compute_elem_cgeom / compute_elem_centroid, mesh_geometry_module.F90:75-127 */
static void elem_centroid(om_mesh *m, int e, double *c) {
  double ce[3] = {0, 0, 0};
  for (int j = 1; j <= m->e_nv[e]; j++)
    for (int d = 0; d < 3; d++) ce[d] = ce[d] + VC(m, m->e_vert[e][j], d);
  for (int d = 0; d < 3; d++) ce[d] = ce[d] / (double)m->e_nv[e];
  double sum_vol = 0.0;
  c[0] = c[1] = c[2] = 0.0;
  for (int j = 1; j <= m->e_nf[e]; j++) {
    int f = m->e_face[e][j];
    const double *cf = &FC(m, f, 0);
    int n = m->f_nv[f];
    for (int k = 1; k <= n; k++) {
      int kp = 1 + (k % n);
      const double *cp = &VC(m, m->f_vert[f][k], 0);
      const double *cpp = &VC(m, m->f_vert[f][kp], 0);
      double a[3], b[3], x[3], fe[3];
      for (int d = 0; d < 3; d++) {
        a[d] = cp[d] - cf[d];
        b[d] = cpp[d] - cf[d];
        fe[d] = cf[d] - ce[d];
      }
      if (m->f_left[f] == e)
        cross3(a, b, x);
      else
        cross3(b, a, x);
      double vol = dot3(x, fe) / 6.0;
      for (int d = 0; d < 3; d++) c[d] = c[d] + vol * (cp[d] + cpp[d] + cf[d] + ce[d]) / 4.0;
      sum_vol = sum_vol + vol;
    }
  }
  for (int d = 0; d < 3; d++) c[d] = c[d] / sum_vol;
}

/* vertex position in a face and its two face-neighbour midpoints */
static void subface_points(om_mesh *m, int f, int v, double *cp, double *cpm, double *cpp) {
  int n = m->f_nv[f], jj = 1;
  for (int j = 1; j <= n; j++)
    if (m->f_vert[f][j] == v) {
      jj = j;
      break;
    }
  int jp = 1 + (jj % n);
  int jm = 1 + ((jj - 2 + n) % n);
  const double *c = &VC(m, v, 0);
  const double *cm = &VC(m, m->f_vert[f][jm], 0);
  const double *cq = &VC(m, m->f_vert[f][jp], 0);
  for (int d = 0; d < 3; d++) {
    cp[d] = c[d];
    cpm[d] = 0.5 * (cm[d] + c[d]);
    cpp[d] = 0.5 * (cq[d] + c[d]);
  }
}

/* compute_subface_norm, mesh_geometry_module.F90:237-271 */
static void subface_norm(om_mesh *m, int sf, double *nrm) {
  int v = m->sf_vert[sf], f = m->sf_face[sf];
  double cp[3], cpm[3], cpp[3];
  subface_points(m, f, v, cp, cpm, cpp);
  const double *cf = &FC(m, f, 0);
  double a[3], b[3], c[3], x1[3], x2[3];
  for (int d = 0; d < 3; d++) {
    a[d] = cpm[d] - cf[d];
    b[d] = cp[d] - cf[d];
    c[d] = cpp[d] - cf[d];
  }
  cross3(a, b, x1);
  cross3(b, c, x2);
  for (int d = 0; d < 3; d++) nrm[d] = 0.5 * (x1[d] + x2[d]);
}

/* compute_subelem_volume, mesh_geometry_module.F90:129-175 */
static double subelem_volume(om_mesh *m, int se) {
  int e = m->se_elem[se];
  const double *ce = &EC(m, e, 0);
  double vol = 0.0;
  for (int j = 1; j <= m->se_nsf[se]; j++) {
    int sf = m->se_sf[se][j];
    int v = m->sf_vert[sf], f = m->sf_face[sf];
    double cp[3], cpm[3], cpp[3];
    subface_points(m, f, v, cp, cpm, cpp);
    const double *cf = &FC(m, f, 0);
    double a[3], b[3], c[3], fe[3], x1[3], x2[3];
    for (int d = 0; d < 3; d++) {
      a[d] = cpm[d] - cf[d];
      b[d] = cp[d] - cf[d];
      c[d] = cpp[d] - cf[d];
      fe[d] = cf[d] - ce[d];
    }
    if (m->f_left[f] == e) {
      cross3(a, b, x1);
      cross3(b, c, x2);
    } else {
      cross3(c, b, x1);
      cross3(b, a, x2);
    }
    vol = vol + dot3(x1, fe) / 6.0 + dot3(x2, fe) / 6.0;
  }
  return vol;
}

/* check_mesh, mesh_geometry_module.F90:290-411 (absolute tolerance 1e-14) */
static int check_mesh(om_mesh *m) {
  int ok = 1;
  for (int i = 1; i <= m->ne; i++) {
    double vol = 0.0, nt[3] = {0, 0, 0}, ns[3] = {0, 0, 0};
    for (int j = 1; j <= m->e_nf[i]; j++) {
      int f = m->e_face[i][j];
      double nr[3];
      for (int d = 0; d < 3; d++) nr[d] = (m->f_left[f] == i) ? FN(m, f, d) : -FN(m, f, d);
      double an[3];
      for (int d = 0; d < 3; d++) an[d] = m->farea[f] * nr[d];
      vol = vol + dot3(&FC(m, f, 0), an);
    }
    vol = vol / 3.0;
    ok = ok && (fabs(m->evol[i] - vol) < 1e-14);
    for (int j = 1; j <= m->e_nf[i]; j++) {
      int f = m->e_face[i][j];
      for (int d = 0; d < 3; d++) {
        double nr = (m->f_left[f] == i) ? FN(m, f, d) : -FN(m, f, d);
        nt[d] = nt[d] + m->farea[f] * nr;
      }
    }
    ok = ok && (nrm3(nt) < 1e-14);
    for (int j = 1; j <= m->e_nf[i]; j++) {
      int f = m->e_face[i][j];
      for (int k = 1; k <= m->f_nsf[f]; k++) {
        int sf = m->f_sf[f][k];
        for (int d = 0; d < 3; d++) {
          double nr = (m->f_left[f] == i) ? SN(m, sf, d) : -SN(m, sf, d);
          ns[d] = ns[d] + m->sf_area[sf] * nr;
        }
      }
    }
    ok = ok && (nrm3(ns) < 1e-14);
  }
  for (int i = 1; i <= m->nse; i++) ok = ok && (m->se_vol[i] > 0.0);
  return ok;
}

/* compute_geometry_mesh, mesh_geometry_module.F90:16-57 */
static int compute_geometry_mesh(om_mesh *m, int b2d) {
  for (int i = 1; i <= m->nf; i++) face_centroid(m, i, &FC(m, i, 0));
  m->ecoord = dalloc(m, 3L * (m->ne + 1));
  for (int i = 1; i <= m->ne; i++) elem_centroid(m, i, &EC(m, i, 0));
  for (int i = 1; i <= m->nsf; i++) {
    double n[3];
    subface_norm(m, i, n);
    double a = nrm3(n);
    m->sf_area[i] = a;
    for (int d = 0; d < 3; d++) SN(m, i, d) = n[d] / a;
  }
  for (int i = 1; i <= m->nf; i++) {
    double acc[3] = {0, 0, 0};
    for (int j = 1; j <= m->f_nsf[i]; j++) {
      double n[3];
      subface_norm(m, m->f_sf[i][j], n);
      for (int d = 0; d < 3; d++) acc[d] = acc[d] + n[d];
    }
    double a = nrm3(acc);
    m->farea[i] = a;
    for (int d = 0; d < 3; d++) FN(m, i, d) = acc[d] / a;
  }
  for (int i = 1; i <= m->nse; i++) m->se_vol[i] = subelem_volume(m, i);
  m->evol = dalloc(m, m->ne);
  for (int i = 1; i <= m->ne; i++) {
    double v = 0.0;
    for (int j = 1; j <= m->e_nse[i]; j++) v = v + m->se_vol[m->e_se[i][j]];
    m->evol[i] = v;
  }
  int ok = check_mesh(m);
  find_if_vert_is_bound(m, b2d);
  m->vvol = dalloc(m, m->nv);
  for (int i = 1; i <= m->nv; i++) {
    double v = 0.0;
    for (int j = 1; j <= m->v_nsen[i]; j++) v = v + m->se_vol[m->v_se[i][j]];
    m->vvol[i] = v;
  }
  return ok;
}

/* --------------------------------------------------------- entry point */
om_mesh *om_mesh_build_p(int nv, const double *coord, int ne, const int *kind, const int *evert,
                         int nbf, const int *bf_nv, const int *bf_vert, const int *bf_tag, int b2d,
                         int periodic, int *status) {
  *status = OM_OK;
  for (int i = 0; i < ne; i++)
    if (n_vert_kind(kind[i]) == 0) {
      *status = OM_ERR_BAD_INPUT;
      return NULL;
    }
  om_mesh *m = (om_mesh *)calloc(1, sizeof(om_mesh));
  read_elements(m, nv, coord, ne, kind, evert, nbf, bf_nv, bf_vert, bf_tag);
  build_node_elem_neigh(m);
  renumber_vert_mesh(m);
  renumber_elem_mesh(m);
  build_faces(m);
  build_node_face_connectivity(m);
  renumber_faces_mesh(m);
  build_sub_faces(m);
  build_sub_elems(m);
  find_sub_elems_around_sub_face(m);
  build_node_sub_face_connectivity(m);
  compute_elem_tag(m);
  m->nmax_sfn = 0;
  for (int i = 1; i <= m->nv; i++)
    if (m->v_nsfn[i] > m->nmax_sfn) m->nmax_sfn = m->v_nsfn[i];
  compute_neigh_by_vert(m);
  if (periodic) {
    if (!stitch_periodic_faces(m) || !build_periodic_stencil(m)) {
      *status = OM_ERR_BAD_INPUT;
      om_mesh_free(m);
      return NULL;
    }
    /* work arrays of the sweeps are sized by the extended lists */
    for (int i = 1; i <= m->nv; i++)
      if (m->v_nsfn[i] > m->nmax_sfn) m->nmax_sfn = m->v_nsfn[i];
  }
  if (!compute_geometry_mesh(m, b2d)) *status = OM_ERR_MESH_GEOMETRY;
  return m;
}

om_mesh *om_mesh_build(int nv, const double *coord, int ne, const int *kind, const int *evert,
                       int nbf, const int *bf_nv, const int *bf_vert, const int *bf_tag, int b2d,
                       int *status) {
  return om_mesh_build_p(nv, coord, ne, kind, evert, nbf, bf_nv, bf_vert, bf_tag, b2d, 0, status);
}

/* counts: nv, ne, nf, nsf, nse, nmax_sfn */
void om_mesh_counts(const om_mesh *m, int *out) {
  out[0] = m->nv;
  out[1] = m->ne;
  out[2] = m->nf;
  out[3] = m->nsf;
  out[4] = m->nse;
  out[5] = m->nmax_sfn;
}

/* Flat exports (1-based ids as stored) for comparisons in tests.
 * Returns the number of values written, or -1 for an unknown name. */
long om_mesh_get_i(const om_mesh *m, const char *name, int *out) {
  long n = 0;
#define FLAT1(arr, cnt) \
  do {                  \
    for (int i = 1; i <= (cnt); i++) out[n++] = (arr)[i]; \
  } while (0)
#define LIST(cntarr, lists, cnt)                                 \
  do {                                                           \
    for (int i = 1; i <= (cnt); i++)                             \
      for (int j = 1; j <= (cntarr)[i]; j++) out[n++] = (lists)[i][j]; \
  } while (0)
  if (!strcmp(name, "vert_nto")) FLAT1(m->vert_nto, m->nv);
  else if (!strcmp(name, "elem_nto")) FLAT1(m->elem_nto, m->ne);
  else if (!strcmp(name, "vert_is_bound")) FLAT1(m->vbound, m->nv);
  else if (!strcmp(name, "vert_n_sub_faces")) FLAT1(m->v_nsfn, m->nv);
  else if (!strcmp(name, "vert_n_sub_elems")) FLAT1(m->v_nsen, m->nv);
  else if (!strcmp(name, "vert_n_elems")) FLAT1(m->v_nen, m->nv);
  else if (!strcmp(name, "vert_sub_face")) LIST(m->v_nsfn, m->v_sf, m->nv);
  else if (!strcmp(name, "vert_sub_elem")) LIST(m->v_nsen, m->v_se, m->nv);
  else if (!strcmp(name, "vert_elem")) LIST(m->v_nen, m->v_elem, m->nv);
  else if (!strcmp(name, "elem_n_vert")) FLAT1(m->e_nv, m->ne);
  else if (!strcmp(name, "elem_kind")) FLAT1(m->e_kind, m->ne);
  else if (!strcmp(name, "elem_vert")) LIST(m->e_nv, m->e_vert, m->ne);
  else if (!strcmp(name, "elem_face")) LIST(m->e_nf, m->e_face, m->ne);
  else if (!strcmp(name, "elem_sub_elem")) LIST(m->e_nse, m->e_se, m->ne);
  else if (!strcmp(name, "elem_tag")) FLAT1(m->e_tag, m->ne);
  else if (!strcmp(name, "elem_n_neigh_by_vert")) FLAT1(m->e_nnbv, m->ne);
  else if (!strcmp(name, "elem_neigh_by_vert")) LIST(m->e_nnbv, m->e_nbv, m->ne);
  else if (!strcmp(name, "face_n_vert")) FLAT1(m->f_nv, m->nf);
  else if (!strcmp(name, "face_vert")) LIST(m->f_nv, m->f_vert, m->nf);
  else if (!strcmp(name, "face_left")) FLAT1(m->f_left, m->nf);
  else if (!strcmp(name, "face_right")) FLAT1(m->f_right, m->nf);
  else if (!strcmp(name, "face_sub_face")) LIST(m->f_nsf, m->f_sf, m->nf);
  else if (!strcmp(name, "sub_face_vert")) FLAT1(m->sf_vert, m->nsf);
  else if (!strcmp(name, "sub_face_face")) FLAT1(m->sf_face, m->nsf);
  else if (!strcmp(name, "sub_face_loc")) FLAT1(m->sf_loc, m->nsf);
  else if (!strcmp(name, "sub_face_left_elem")) FLAT1(m->sf_le, m->nsf);
  else if (!strcmp(name, "sub_face_right_elem")) FLAT1(m->sf_re, m->nsf);
  else if (!strcmp(name, "sub_face_left_sub_elem")) FLAT1(m->sf_lse, m->nsf);
  else if (!strcmp(name, "sub_face_right_sub_elem")) FLAT1(m->sf_rse, m->nsf);
  else if (!strcmp(name, "sub_elem_loc")) FLAT1(m->se_loc, m->nse);
  else if (!strcmp(name, "sub_elem_vert")) FLAT1(m->se_vert, m->nse);
  else if (!strcmp(name, "sub_elem_elem")) FLAT1(m->se_elem, m->nse);
  else if (!strcmp(name, "sub_elem_sub_face")) LIST(m->se_nsf, m->se_sf, m->nse);
  else return -1;
#undef FLAT1
#undef LIST
  return n;
}

long om_mesh_get_d(const om_mesh *m, const char *name, double *out) {
  long n = 0;
  if (!strcmp(name, "vert_coord")) {
    for (int i = 1; i <= m->nv; i++)
      for (int d = 0; d < 3; d++) out[n++] = VC(m, i, d);
  } else if (!strcmp(name, "vert_volume")) {
    for (int i = 1; i <= m->nv; i++) out[n++] = m->vvol[i];
  } else if (!strcmp(name, "elem_coord")) {
    for (int i = 1; i <= m->ne; i++)
      for (int d = 0; d < 3; d++) out[n++] = EC(m, i, d);
  } else if (!strcmp(name, "elem_volume")) {
    for (int i = 1; i <= m->ne; i++) out[n++] = m->evol[i];
  } else if (!strcmp(name, "face_coord")) {
    for (int i = 1; i <= m->nf; i++)
      for (int d = 0; d < 3; d++) out[n++] = FC(m, i, d);
  } else if (!strcmp(name, "face_norm")) {
    for (int i = 1; i <= m->nf; i++)
      for (int d = 0; d < 3; d++) out[n++] = FN(m, i, d);
  } else if (!strcmp(name, "face_area")) {
    for (int i = 1; i <= m->nf; i++) out[n++] = m->farea[i];
  } else if (!strcmp(name, "sub_face_norm")) {
    for (int i = 1; i <= m->nsf; i++)
      for (int d = 0; d < 3; d++) out[n++] = SN(m, i, d);
  } else if (!strcmp(name, "sub_face_area")) {
    for (int i = 1; i <= m->nsf; i++) out[n++] = m->sf_area[i];
  } else if (!strcmp(name, "sub_elem_volume")) {
    for (int i = 1; i <= m->nse; i++) out[n++] = m->se_vol[i];
  } else
    return -1;
  return n;
}
