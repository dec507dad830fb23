/* A plain C caller of include/semx.h: links libsemx.so and runs the
 * host-only entry points (mesh construction and numbering, status text,
 * cylindrical coordinates and a field-named system's setup), printing what
 * a test compares.  Built as C and as C++ by
 * tests/test_capi_client.py. */
#include <stdio.h>
#include <stdlib.h>

#include "semx.h"

int main (void)
{
  /* 2 x 1 quads on [0,2]x[0,1], Dirichlet all round except the joined
   * side: element 0 side 1 (right) meets element 1 side 3 (left) */
  const double vx[8] = {0, 1, 1, 0,  1, 2, 2, 1};
  const double vy[8] = {0, 0, 1, 1,  0, 0, 1, 1};
  const int nbr[8]   = {-1, 7, -1, -1,  -1, -1, -1, 1};
  const int skind[8] = {1, 0, 1, 1,  1, 1, 1, 0};
  const int np = 4, nel = 2, next = 4 * (np - 1);
  semx_mesh m = NULL, m2 = NULL;
  int info[10], st, i;
  int btog[2 * 12], bmask[2 * 12], emask[2];

  st = semx_mesh_create (np, nel, vx, vy, nbr, skind, -1, 0.0, 1.0, &m);
  if (st) { printf ("create %d %s\n", st, semx_error_string (st)); return 1; }
  semx_mesh_info (m, info);
  printf ("info");
  for (i = 0; i < 10; i++) printf (" %d", info[i]);
  printf ("\n");
  semx_mesh_numbering (m, btog, bmask, emask);
  /* the same mesh with the numbering handed back */
  st = semx_mesh_create_numbered (np, nel, vx, vy, nbr, skind, -1, 0.0, 1.0, btog, bmask,
                                  info[4], info[5], &m2);
  printf ("numbered %d\n", st);
  /* a numbering that joins two distinct nodes is refused */
  btog[0] = btog[next - 1] == btog[0] ? btog[1] : btog[next - 1];
  {
    semx_mesh m3 = NULL;
    const int st3 = semx_mesh_create_numbered (np, nel, vx, vy, nbr, skind, -1, 0.0, 1.0,
                                               btog, bmask, info[4], info[5], &m3);
    printf ("refused %d %s\n", st3, semx_error_string (st3));
    if (m3) semx_mesh_destroy (m3);
  }
  printf ("bad-arg %d\n", semx_mesh_create (2, 1, vx, vy, nbr, skind, -1, 0.0, 1.0, &m2));
  /* cylindrical coordinates (y radial, 0 <= y <= 1) and a system for the
   * coupled radial field v~ (host setup only) */
  {
    semx_mesh mc = NULL;
    semx_helm hv = NULL;
    double hbb[12 * 12], hbi[12 * 4], hii[4 * 4];
    int stc = semx_mesh_create (np, nel, vx, vy, nbr, skind, -1, 0.0, 1.0, &mc);
    if (!stc) stc = semx_mesh_set_cylindrical (mc);
    printf ("cyl %d\n", stc);
    stc = semx_helm_create_field (mc, 1.0, 1.0, 0, 2, 4, SEMX_DIRECT, 'v', &hv);
    printf ("field-v %d\n", stc);
    if (!stc && !semx_helm_element (hv, 1, 0, hbb, hbi, hii))
      printf ("hbb00-positive %d\n", hbb[0] > 0.0);
    printf ("cyl-after-system %d\n", semx_mesh_set_cylindrical (mc));
    printf ("field-x %d\n", semx_helm_create_field (mc, 1.0, 1.0, 0, 2, 4, SEMX_DIRECT, 'x',
                                                     &hv));
    if (hv) semx_helm_destroy (hv);
    semx_mesh_destroy (mc);
  }
  semx_mesh_destroy (m);
  return 0;
}
