/*
 * util.h -- small helpers and binary I/O (drop-in for the reference's include/util.h:80-82).
 * ReadData returns -1 (cannot open), -2 (size mismatch) or -3 (short read); WriteData -1 / -3.
 */
#ifndef UTIL_H
#define UTIL_H

#include <math.h>
#include <stdbool.h>
#include <stddef.h>

static inline double square(const double x) { return x * x; }
static inline double maxf(const double x, const double y) { return x >= y ? x : y; }
static inline double minf(const double x, const double y) { return x <= y ? x : y; }

int ReadData(const char *filename, void *data, const size_t size, const size_t n);
int WriteData(const char *filename, const void *data, const size_t size, const size_t n, const bool append);

/* qbf extension: message of the last error recorded by the library (empty if none) */
const char *qbf_last_error(void);

#endif
