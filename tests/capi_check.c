/* capi_check.c -- a C program against include/d2hip.h (no Python): builds
 * the reference's pyramid and plane meshes (test_comparison.cpp:93-105,
 * primitive_geo.cpp) into trees with d2_tree_build, uploads them with two
 * bodies placed as test_comparison.cpp:90-128 places them, and checks through
 * the C ABI: the discrete narrow phase (8 raw contacts, 1 after
 * filterContacts), the plugin point's bucket comparison (8 hits, pair_used 8),
 * the standalone filter (1 kept), and a CCD run of the plane descending onto
 * the pyramid (test_continuous_comparison.cpp:222-263: first contact at toc
 * 0.9).  Compiled with gcc as C99 (the header must stay C); exits 0 on
 * success, prints the failing check otherwise.  With argument "link" it only
 * resolves every entry point and exits (no GPU needed). */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "d2hip.h"

#define CHECK(cond, ...)                    \
  do {                                      \
    if (!(cond)) {                          \
      fprintf(stderr, "capi_check: ");      \
      fprintf(stderr, __VA_ARGS__);         \
      fprintf(stderr, "\n");                \
      return 1;                             \
    }                                       \
  } while (0)

static void body(double* r, double z, double vz, int is_static, double fz) {
  memset(r, 0, 40 * sizeof(double));
  r[0] = 1.0;  /* current: identity rotation */
  r[6] = z;
  r[9] = vz;
  r[14] = 1.0; /* future */
  r[20] = fz;
  r[23] = vz;
  r[28] = r[32] = r[36] = 1.0; /* inverse inertia */
  r[37] = 0.01;                /* geo_eps */
  r[38] = is_static ? 1.0 : 0.0;
  r[39] = 1.0;                 /* last_time_step_size */
}

int main(int argc, char** argv) {
  if (argc > 1 && strcmp(argv[1], "link") == 0) {
    /* every entry point resolves (the link already proved it; touch them) */
    typedef void (*fn_t)(void);
    fn_t fns[] = {(fn_t)d2_ctx_create, (fn_t)d2_narrow_run, (fn_t)d2_ccd_run, (fn_t)d2_step_run,
                  (fn_t)d2_find_contacts_bucket, (fn_t)d2_find_contacts_bucket_noalloc,
                  (fn_t)d2_filter_contacts, (fn_t)d2_broad_phase,
                  (fn_t)d2_clusters_build, (fn_t)d2_tree_build, (fn_t)d2_touch_stats};
    int linked = 0;
    for (size_t i = 0; i < sizeof(fns) / sizeof(fns[0]); ++i) linked += fns[i] != NULL;
    printf("capi_check: %d entry points linked\n", linked);
    return 0;
  }
  static const double pyramid_v[] = {1, 1, 0, -1, 1, 0, -1, -1, 0, 1, -1, 0, 0, 0, 0.1};
  static const int32_t pyramid_f[] = {0, 1, 4, 1, 2, 4, 2, 3, 4, 3, 0, 4};
  static const double plane_v[] = {0.5, 0.5, 0, 0.5, -0.5, 0, -0.5, 0.5, 0, -0.5, -0.5, 0};
  static const int32_t plane_f[] = {0, 2, 1, 1, 2, 3};
  uint64_t rng[2];
  int32_t nodes[2][8 * 8];
  double buckets[2][8 * 88];
  int n_nodes[2];
  CHECK(d2_pcg_seed(42, 54, rng) == D2_OK, "pcg seed");
  CHECK(d2_tree_build(pyramid_v, 5, pyramid_f, 4, rng, 8, nodes[0], buckets[0], &n_nodes[0]) == D2_OK, "pyramid tree");
  CHECK(d2_tree_build(plane_v, 4, plane_f, 2, rng, 8, nodes[1], buckets[1], &n_nodes[1]) == D2_OK, "plane tree");
  CHECK(n_nodes[0] == 1 && n_nodes[1] == 1, "single-leaf trees (%d, %d nodes)", n_nodes[0], n_nodes[1]);
  int32_t all_nodes[16];
  double all_buckets[176];
  memcpy(all_nodes, nodes[0], 8 * sizeof(int32_t));
  memcpy(all_nodes + 8, nodes[1], 8 * sizeof(int32_t));
  memcpy(all_buckets, buckets[0], 88 * sizeof(double));
  memcpy(all_buckets + 88, buckets[1], 88 * sizeof(double));
  const int32_t tree_off[] = {0, 1, 2};

  d2_ctx* c = NULL;
  int rc = d2_ctx_create(0, D2_FP64, &c);
  CHECK(rc == D2_OK, "d2_ctx_create: %d (is a GPU visible?)", rc);
  CHECK(d2_trees_upload(c, 2, tree_off, all_nodes, all_buckets) == D2_OK, "trees: %s", d2_last_error(c));
  double rec[80];
  body(rec, 1.0, 0.0, 1, 1.0);       /* pyramid, static, future at z = 1 */
  body(rec + 40, 1.105, 0.0, 0, 1.105); /* plane 0.005 above the apex */
  const int32_t tob[] = {0, 1};
  CHECK(d2_bodies_upload(c, 2, tob, rec) == D2_OK, "bodies: %s", d2_last_error(c));

  /* compareTreesFull + filterContacts (test_comparison.cpp:90-128) */
  const int32_t pair[] = {0, 1};
  int64_t n = 0;
  CHECK(d2_narrow(c, 1, pair, 0, &n) == D2_OK && n == 8, "narrow: %lld contacts (%s)", (long long)n, d2_last_error(c));
  CHECK(d2_narrow(c, 1, pair, 1, &n) == D2_OK && n == 1, "narrow + filter: %lld contacts", (long long)n);

  /* the plugin point's bucket comparison over the root node pair */
  const int32_t np[] = {0, 0, 0, 0};
  int32_t used[2], ints[16 * 8];
  double reals[16 * 11];
  CHECK(d2_find_contacts_bucket(c, 0, 1, 2, np, used, 16, ints, reals, &n) == D2_OK, "bucket: %s", d2_last_error(c));
  CHECK(n == 16 && used[0] == 8 && used[1] == 16, "bucket: %lld hits, pair_used %d %d", (long long)n, used[0], used[1]);
  CHECK(d2_find_contacts_bucket(c, 0, 1, 1, np, used, 4, ints, reals, &n) == D2_ERR_CAPACITY && n == 8,
        "bucket over max_hits must report D2_ERR_CAPACITY with the count");

  /* the standalone filter over the 8 bucket hits */
  int32_t order[8], kept[8];
  int n_kept = 0;
  CHECK(d2_find_contacts_bucket(c, 0, 1, 1, np, used, 8, ints, reals, &n) == D2_OK, "bucket: %s", d2_last_error(c));
  CHECK(d2_filter_contacts(c, 8, reals, 0.0, order, kept, &n_kept) == D2_OK && n_kept == 1, "filter: %d kept", n_kept);

  /* CCD: the plane from z = 2 at -1 m/s over [0, 1] meets the apex (z = 1.1) at 0.9 */
  body(rec + 40, 2.0, -1.0, 0, 2.0);
  CHECK(d2_bodies_upload(c, 2, tob, rec) == D2_OK, "bodies: %s", d2_last_error(c));
  const double max_time = 1.0;
  CHECK(d2_ccd(c, 1, pair, &max_time, &n) == D2_OK && n > 0, "ccd: %lld contacts", (long long)n);
  double ccd_reals[64 * 11];
  int32_t ccd_ints[64 * 8], off[2];
  CHECK(n <= 64 && d2_result_contacts(c, off, ccd_ints, ccd_reals) == D2_OK, "ccd results");
  double toc = 1.0;
  for (int64_t k = 0; k < n; ++k) toc = ccd_reals[11 * k + 10] < toc ? ccd_reals[11 * k + 10] : toc;
  CHECK(fabs(toc - 0.9) < 1e-5, "ccd: first toc %.9f, want 0.9", toc);

  CHECK(d2_ctx_destroy(c) == D2_OK, "destroy");
  printf("capi_check: ok\n");
  return 0;
}
