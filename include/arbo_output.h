/* arbo_output.h — the reference's output files from engine series.
 *
 * arbo_write_ndata replaces writeIncidence (writeoutput.cpp:13-98): the
 * nData<mcmcRun>.csv time series of one simulation run, created with its
 * header by run 0 and appended to by later runs.  Columns: t, popSize,
 * incidence_DENV1..S, prevalence_DENV1..S,
 * prevalence_DENV1..S_random_community, simRun; one row per day 0..rows-1,
 * each written after a newline (the file has no trailing newline).
 *
 * arbo_write_nspatial replaces writeSpatial (writeoutput.cpp:194-241): the
 * nSpatial<mcmcRun>.csv file of infectious humans per community and
 * serotype.  Header t, subPop, prevalence.DENV1..S, simRun; then one line
 * per day and community (t, subPop, one value per serotype), where, as in
 * the reference, the simRun column is printed only after the last
 * community of each day.
 *
 * Conventions as include/arbo_c.h.
 */
#ifndef ARBO_OUTPUT_H
#define ARBO_OUTPUT_H
#include <stdint.h>

#include "arbo_c.h"

#ifdef __cplusplus
extern "C" {
#endif

/* path: the file (the caller composes odir/nData<mcmcRun>.csv); series: one
 * simulation's series as read by arbo_read_series, `rows` rows of S
 * serotypes; sim_run 0 creates the file, any other appends. */
ARBO_API int arbo_write_ndata(const char* path, const ArboSeries* series, uint32_t rows, uint32_t serotypes,
                              uint32_t sim_run);

/* spatial: rows * serotypes * communities words as read by
 * arbo_read_spatial; sim_run 0 creates the file, any other appends. */
ARBO_API int arbo_write_nspatial(const char* path, const uint32_t* spatial, uint32_t rows, uint32_t serotypes,
                                 uint32_t communities, uint32_t sim_run);

/* The partial human census of writeCensus (writeoutput.cpp:100-148, the
 * nAlive file): header "age,history1..S,simRun", then one line for each of
 * the first min(5000, n_humans) humans.  age: each human's age in days
 * after the last simulated day, (days done - 1) - h_birth; history: the
 * S * n_humans words of arbo_read_history ([s * n_humans + i]).  sim_run 0
 * creates the file, any other appends. */
ARBO_API int arbo_write_nalive(const char* path, const uint32_t* age, const int32_t* history, uint32_t n_humans,
                               uint32_t serotypes, uint32_t sim_run);

/* The age-of-infection file of writeAOI (writeoutput.cpp:152-191, nAOI):
 * header "first,second,third,fourth,simRun", then ARBO_AOI_RECORD lines of
 * the four rings of arbo_read_aoi (aoi[k * ARBO_AOI_RECORD + r]).  sim_run 0
 * creates the file, any other appends. */
ARBO_API int arbo_write_naoi(const char* path, const uint32_t* aoi, uint32_t sim_run);

#ifdef __cplusplus
}
#endif
#endif /* ARBO_OUTPUT_H */
