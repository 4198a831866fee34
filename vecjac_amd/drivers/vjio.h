/* vjio.h - file and timing helpers of the command-line drivers.
 *
 * The drivers read and write the reference's raw formats (mtxio.c:69-213):
 * BaseName.EXT holds a matrix column by column with no header (m doubles,
 * or m interleaved double-complex pairs, per column); singular values are
 * written as x86 extended long doubles (the reference's `wide` under
 * USE_EXTENDED, common.h:86), 16 bytes each.  Timings use
 * CLOCK_MONOTONIC_RAW, as the reference's timer does.
 */
#ifndef VJIO_H
#define VJIO_H

#define _GNU_SOURCE
#include <fcntl.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include "vecjac.h"

static inline int vj_open(const char *bn, const char *ex, int wr)
{
  const size_t l = strlen(bn) + strlen(ex) + 2u;
  char *fn = (char *)malloc(l);
  if (!fn)
    return -1;
  (void)snprintf(fn, l, "%s.%s", bn, ex);
  const int fd = wr ? open(fn, O_WRONLY | O_CREAT | O_TRUNC, S_IRUSR | S_IWUSR | S_IRGRP | S_IROTH) : open(fn, O_RDONLY);
  free(fn);
  return fd;
}

/* rows 0..m-1 of n columns of `w` bytes per element, ld elements apart in
 * memory and contiguous in the file */
static inline int vj_read_cols(int fd, uint64_t m, uint64_t n, void *A, uint64_t ld, size_t w)
{
  for (uint64_t j = 0; j < n; ++j) {
    const size_t sz = (size_t)m * w;
    if ((ssize_t)sz != pread(fd, (char *)A + (size_t)j * ld * w, sz, (off_t)(j * sz)))
      return 1;
  }
  return 0;
}

static inline int vj_write_cols(int fd, uint64_t m, uint64_t n, const void *A, uint64_t ld, size_t w)
{
  for (uint64_t j = 0; j < n; ++j) {
    const size_t sz = (size_t)m * w;
    if ((ssize_t)sz != pwrite(fd, (const char *)A + (size_t)j * ld * w, sz, (off_t)(j * sz)))
      return 1;
  }
  return 0;
}

static inline long double vj_now(void)
{
  struct timespec t;
  (void)clock_gettime(CLOCK_MONOTONIC_RAW, &t);
  return (long double)t.tv_sec + (long double)t.tv_nsec * 1e-9L;
}

static inline void *vj_alloc(size_t bytes)
{
  void *p = NULL;
  return posix_memalign(&p, 64u, bytes ? ((bytes + 63u) & ~(size_t)63u) : 64u) ? NULL : p;
}

/* sigma_j = fS_j 2^eS_j in extended precision, written as n long doubles */
static inline int vj_write_sigma(const char *bn, const char *ex, uint64_t n, const double *eS, const double *fS)
{
  long double *s = (long double *)vj_alloc(n * sizeof(long double));
  if (!s)
    return 1;
  for (uint64_t i = 0; i < n; ++i)
    s[i] = (eS[i] != 0.0) ? scalbnl((long double)fS[i], isfinite(eS[i]) ? (int)eS[i] : (eS[i] < 0 ? -100000 : 100000))
                          : (long double)fS[i];
  const int fd = vj_open(bn, ex, 1);
  int r = (fd < 0);
  if (!r)
    r = ((ssize_t)(n * sizeof(long double)) != write(fd, s, n * sizeof(long double)));
  if (fd >= 0)
    r |= close(fd);
  free(s);
  return r;
}

#endif /* !VJIO_H */
