/*
 * scc_dropin.h -- C++ drop-in for the reference's library interface.
 *
 * Declares, with C++ linkage and the reference's exact parameter lists, the
 * five drivers of include/scc.h:49-53 and the loader of include/load.h:6, as
 * exported by libscc_amd.so.  A program written against the reference headers
 * (src/scc.cu:74, 78-102 calls these) links against libscc_amd.so unchanged:
 * the mangled names are the reference's.  The standard headers that the
 * reference's scc.h brought in (scc.h:4-6: <iostream>, <fstream>, <stdint.h>)
 * come with this one, plus <cstdio>/<cstdlib>, which callers of the reference
 * used through them (printf, atoi).
 *
 * Behaviour at this boundary:
 *  - Graph arrays are the reference's host CSR (1-based vertices, dummy
 *    vertex 0, Fr/Br with V+2 entries, CSize = E, RSize = V).
 *  - Each driver runs on device 0 and prints the reference's stdout lines
 *    (scc_WCC.cu:948, 966-967; scc_Coloring.cu:669-670; scc_WCC.cu:1576-1577).
 *  - On any failure a driver throws a `const char *` message, as the
 *    reference throws on device out-of-memory (scc_WCC.cu:618-623) and its
 *    main catches (scc.cu:104-108); the reference's exit() from inside the
 *    library on CUDA errors (scc.h:12-20) is not reproduced.
 *  - loadFullGraph allocates with new[] (the caller's delete[] of
 *    scc.cu:110-113 is correct).  On a missing or malformed file it prints
 *    "Error Reading graph file" (load.cpp:26) and sets every output to 0 /
 *    NULL.  *oPr is read from the file named by the SCC_PARTITION_FILE
 *    environment variable (one 0-based partition id per line, V lines, into
 *    Pr[1..V]) when it is set, else NULL; the reference read a hard-coded
 *    path (load.cpp:117-130).  detectSCC throws when Pr is NULL.
 */
#ifndef SCC_DROPIN_H
#define SCC_DROPIN_H

#ifndef __cplusplus
#error "scc_dropin.h is the C++ interface; C and FFI callers use scc_amd.h (scc_vHong, ...)"
#endif

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdint.h>

/* include/scc.h:49-53 */
void wSlota(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
            bool t1, bool t2, int warpSize);
void vSlota(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
            bool t1, bool t2);
void wHong(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
           bool t1, bool t2, int warpSize);
void vHong(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
           bool t1, bool t2);
void detectSCC(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc,
               uint32_t *Br, uint32_t *Pr, bool t1, bool t2);

/* include/load.h:6 */
void loadFullGraph(const char *filename, uint32_t *oCSize, uint32_t *oRSize, uint32_t **oFc,
                   uint32_t **oFr, uint32_t **oBc, uint32_t **oBr, uint32_t **oPr);

#endif
