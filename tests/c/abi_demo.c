/* A C (not C++) caller of include/octree_pipeline.h: builds the octree of
 * 4096 uniform points with one ocp_build and checks the counts against the
 * ones passed on the command line. Without a GPU it stops after ocp_create
 * and exits 3. Compiled and linked by tests/test_abi.py; run by the GPU
 * test there. */
#include <stdio.h>
#include <stdlib.h>

#include "octree_pipeline.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: abi_demo <n_unique> <n_oct>\n");
    return 2;
  }
  const int32_t want_u = atoi(argv[1]), want_k = atoi(argv[2]);
  ocp_ctx ctx;
  if (ocp_create(&ctx, 0, NULL) != OCP_OK) {
    fprintf(stderr, "no device\n");
    return 3;
  }
  enum { N = 4096 };
  static float pts[N * 4];
  unsigned s = 12345u;
  for (int i = 0; i < N * 4; ++i) {
    s = s * 1664525u + 1013904223u;
    pts[i] = (i % 4 == 3) ? 0.0f : (float)(s >> 8) / 16777216.0f * 1000.0f;
  }
  void* d_pts = NULL;
  int rc = ocp_malloc(ctx, &d_pts, sizeof(pts));
  if (!rc) rc = ocp_memcpy_h2d(ctx, d_pts, pts, sizeof(pts));
  if (!rc) rc = ocp_build(ctx, (const float*)d_pts, 0, N, 0.0f, 1024.0f);
  int32_t u = -1, k = -1;
  if (!rc) rc = ocp_get_counts(ctx, &u, &k);
  if (rc) fprintf(stderr, "error %d: %s\n", rc, ocp_last_error(ctx));
  ocp_free(ctx, d_pts);
  ocp_destroy(ctx);
  printf("n_unique %d n_oct %d\n", u, k);
  return (!rc && u == want_u && k == want_k) ? 0 : 1;
}
