/* arbo_data.h — the reference's input side: the daily climate and weekly
 * dengue files it reads (ClimateData.cpp:53-126, DengueData.cpp:49-112) and
 * the degree-heterogeneous community sizes of calcCumulativeComSize
 * (network.cpp:116-164) with the per-individual community draw of
 * demographicintial.cu.
 *
 * Files: a header line, then one line per day (climate) or week (dengue).
 * Columns are found by their header name, with surrounding double quotes
 * removed; entry j of a data line belongs to column j mod (header columns),
 * as in the reference; every entry of a named column is parsed as a float
 * (std::stof: leading blanks, trailing text ignored).  Each reader returns
 * the number of data lines in *rows and fills at most `cap` of them; call
 * with cap 0 to size the arrays first.  Errors: the file cannot be opened, a
 * named column is missing, or an entry of a named column is not a number. */
#ifndef ARBO_DATA_H
#define ARBO_DATA_H

#include <stdint.h>

#include "arbo_c.h"

#ifdef __cplusplus
extern "C" {
#endif

/* climate_smooth.csv: columns "T" (deg C), "H" (relative humidity), "P"
 * (precipitation); NULL arrays are skipped.  The arrays feed ArboClimate.
 * Both readers take columns by header name, strip double quotes and accept
 * LF or CRLF line endings; a cell that does not start with a number is an
 * error. */
ARBO_API int arbo_read_climate_csv(const char* path, float* temperature, float* humidity, float* precipitation,
                                   uint32_t cap, uint32_t* rows);

/* dengue.csv: column "incidence", one row per week (the MCMC data). */
ARBO_API int arbo_read_dengue_csv(const char* path, float* incidence, uint32_t cap, uint32_t* rows);

/* Community sizes of a grid_w x grid_h lattice (neighbourhood 4 or 8,
 * include_self as in ArboParams): community c is drawn with probability
 * proportional to degree(c)^hetero, degree = its dispersal-row length
 * (network.cpp:119-164: nHeteroPopSize / mHeteroPopSize, 0 = equal sizes).
 * Every community first receives one individual (the engine needs at least
 * one human per community); each of the other total - C individuals draws
 * its community independently, a 24-bit Philox uniform keyed by (seed,
 * individual) against the cumulative 24-bit thresholds.  out[c], c =
 * y * grid_w + x.  Errors: total < C or above 2^32 - 1, bad lattice, hetero
 * not finite. */
ARBO_API int arbo_community_sizes(uint32_t grid_w, uint32_t grid_h, uint32_t neighbourhood, uint32_t include_self,
                                  uint64_t total, float hetero, uint64_t seed, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* ARBO_DATA_H */
