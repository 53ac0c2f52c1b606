/*
 * fwh_mesh.h - plain-array description of a decomposed full-wave model.
 *
 * This is the data the reference keeps in its PartMesh / DomMesh structs
 * (lib/std/types.h:432-484 DomMesh, :515-563 PartMesh, :292-299 Inbc,
 * :302-305 OPinfo), flattened into C arrays so that it can cross a C ABI
 * without any of the reference's types.  Subdomains are stored part-major
 * (all subdomains of part 0, then part 1, ...) in the reference's loop order
 * (lib/hddm/hddm_solver_s.h:23-43).
 *
 * Conventions kept from the reference:
 *  - element connectivity: 10 ids per tetrahedron, 4 local vertex ids then
 *    6 local edge (mid-node) ids; edge k joins tet vertices
 *    (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) for k = 0..5, i.e. the order after
 *    the reference's 8<->9 mid-node swap (lib/fem/make_MESH.c:573-584,
 *    table tet_n in lib/fem/tetrahedral.h:19-24);
 *  - the six-times volume of every tetrahedron is positive;
 *  - a subdomain's Dirichlet list is [essential edges | interface edges]
 *    (lib/hddm/make_PartMesh.c:1082-1163); interface values are renewed
 *    in the tail (lib/hddm/hddm_solver_sp.h:58-69);
 *  - interface edge i of a subdomain maps to index ln_part of its part's
 *    interface vectors; ln_part < outfree are copies owned by a lower part,
 *    outfree <= ln_part < outfree+infree are owned by this part;
 *  - op lists: for parts p != q, op_ibid[op_off[p*npart+q] ..] lists the
 *    ln_part indices (in part p) of the edges p shares with q, in an order
 *    matched element-by-element with the (q,p) list
 *    (lib/hddm/hddm_solver_s.h:404-424).
 *
 * Material values are resolved per element (the reference looks them up by
 * flag, lib/fem/make_elem_em.c:98-115): absolute permittivity eps, and
 * conductivity sigma.  src[e] != 0 marks an element that carries the vertex
 * current source; joe holds, per local vertex, the complex 3-vector the
 * reference calls Joe = -(Jo + i Jor) (lib/fem/make_elem_em.c:291-296).
 * pml, when not NULL, holds per element the complex coordinate stretch
 * factors (sx, sy, sz) as 6 doubles (re, im interleaved); stretch 1 is the
 * reference element.
 * dgv, when not NULL, selects the delta-gap source of a volt.dat with dg > 0
 * (lib/fem/make_elem_em.c:320-396): per local vertex the gap voltage dV of
 * the last gap box containing the vertex, 0 outside every box.  The element
 * load is then sigma*omega * sum_j (grad lambda_j, N_i) * i dV_j and the
 * current source (src/joe) is not used, as in the reference.
 */
#ifndef FWH_MESH_H
#define FWH_MESH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fwh_mesh_desc {
  int npart;              /* number of parts                                */
  int nsub;               /* total number of subdomains (part-major)        */
  double omega;           /* angular frequency (AntennaOmega)               */
  const int *sub_part;    /* [nsub] part of each subdomain                  */
  const int *el_off;      /* [nsub+1] element offsets                       */
  const int *nop;         /* [nel*10] local vertex ids, local edge ids      */
  const double *eps;      /* [nel] absolute permittivity                    */
  const double *sigma;    /* [nel] conductivity                             */
  const int *src;         /* [nel] current-source element flag              */
  const double *pml;      /* [nel*6] or NULL: complex (sx,sy,sz)            */
  const int *ap_off;      /* [nsub+1] vertex offsets                        */
  const double *crd;      /* [nap*3] vertex coordinates                     */
  const double *joe;      /* [nap*6] complex source vector per vertex       */
  const int *nmp;         /* [nsub] edge DOFs per subdomain                 */
  const int *bc_off;      /* [nsub+1] essential-BC offsets                  */
  const int *ibc;         /* essential-BC local edge ids                    */
  const int *ib_off;      /* [nsub+1] interface offsets                     */
  const int *ib_nfid;     /* interface local edge ids                       */
  const int *ib_ln;       /* interface part-level index (ln_part)           */
  const int *infree;      /* [npart] owned interface DOFs                   */
  const int *outfree;     /* [npart] non-owned interface DOFs               */
  const int *op_off;      /* [npart*npart+1] offsets into op_ibid           */
  const int *op_ibid;     /* shared-DOF lists per part pair                 */
  const double *dgv;      /* [nap] or NULL: delta-gap voltage per vertex    */
} fwh_mesh_desc;

#ifdef __cplusplus
}
#endif

#endif /* FWH_MESH_H */
