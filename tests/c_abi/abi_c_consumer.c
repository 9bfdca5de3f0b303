/* A plain C11 consumer of include/gpufft/gpufft.h: the header compiles as C
 * and the exports link from C.  Calls only host-side entry points (no GPU
 * needed); prints one line per check and exits non-zero on a mismatch. */
#include <stdio.h>
#include <string.h>

#include "gpufft/gpufft.h"

static int fails = 0;

static void check(int cond, const char* what) {
  printf("%s %s\n", cond ? "ok" : "FAIL", what);
  if (!cond) ++fails;
}

int main(void) {
  char buf[8192];
  size_t bytes = 0;
  gfft_plan plan = NULL;
  check(gfft_version() == GFFT_VERSION, "gfft_version matches the header");
  check(strlen(gfft_strerror(GFFT_ERR_INVALID_ARG)) > 0, "gfft_strerror");
  check(gfft_query_plan(512, 512, 512, 1, 0, GFFT_F32, buf, sizeof buf) == GFFT_OK &&
            strstr(buf, "\"K\":257") != NULL,
        "gfft_query_plan 512^3 FP32 reports K = 257");
  check(gfft_query_plan(17, 17, 34, 1, 0, GFFT_F32, buf, sizeof buf) == GFFT_ERR_UNSUPPORTED_SIZE,
        "prime factor 17 is rejected");
  check(gfft_query_plan(8, 8, 8, 1, 0, GFFT_F32, buf, 4) == GFFT_ERR_BUFFER_TOO_SMALL, "short text buffer");
  check(gfft_request_bytes(64, 64, 64, GFFT_F64, &bytes) == GFFT_OK && bytes == (size_t)64 * 64 * 66 * 8,
        "gfft_request_bytes 64^3 FP64");
  check(gfft_plan_create(NULL, 8, 8, 8, 1, 0, GFFT_F32) == GFFT_ERR_INVALID_ARG, "null plan pointer");
  check(gfft_plan_destroy(plan) == GFFT_ERR_INVALID_ARG, "destroy of a null plan");
  check(gfft_exec_r2c(NULL, NULL, NULL) == GFFT_ERR_INVALID_ARG, "exec with a null plan");
  return fails ? 1 : 0;
}
