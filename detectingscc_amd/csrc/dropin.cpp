// dropin.cpp -- the reference's C++ library interface (include/scc.h:49-53,
// include/load.h:6) with C++ linkage and the reference's parameter lists,
// over the C ABI of libscc_amd.so.  Failures throw a `const char *` message,
// which the reference's main catches and prints (src/scc.cu:104-108).
#include <cstring>

#pragma GCC visibility push(default)
#include "scc_dropin.h"
#pragma GCC visibility pop

#include "scc_amd.h"

namespace {

void check(scc_status s)
{
    if (s != SCC_OK) throw scc_last_error();
}

uint32_t *copy_new(const uint32_t *p, size_t n)
{
    uint32_t *q = new uint32_t[n ? n : 1];
    if (n) std::memcpy(q, p, n * sizeof(uint32_t));
    return q;
}

}  // namespace

void vHong(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
           bool t1, bool t2)
{
    check(scc_vHong(CSize, RSize, Fc, Fr, Bc, Br, t1, t2));
}

void wHong(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
           bool t1, bool t2, int warpSize)
{
    check(scc_wHong(CSize, RSize, Fc, Fr, Bc, Br, t1, t2, warpSize));
}

void vSlota(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
            bool t1, bool t2)
{
    check(scc_vSlota(CSize, RSize, Fc, Fr, Bc, Br, t1, t2));
}

void wSlota(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc, uint32_t *Br,
            bool t1, bool t2, int warpSize)
{
    check(scc_wSlota(CSize, RSize, Fc, Fr, Bc, Br, t1, t2, warpSize));
}

void detectSCC(uint32_t CSize, uint32_t RSize, uint32_t *Fc, uint32_t *Fr, uint32_t *Bc,
               uint32_t *Br, uint32_t *Pr, bool t1, bool t2)
{
    check(scc_detectSCC(CSize, RSize, Fc, Fr, Bc, Br, Pr, t1, t2));
}

// load.cpp:12-140 without the hard-coded partition path: Pr comes from
// SCC_PARTITION_FILE when it is set.
void loadFullGraph(const char *filename, uint32_t *oCSize, uint32_t *oRSize, uint32_t **oFc,
                   uint32_t **oFr, uint32_t **oBc, uint32_t **oBr, uint32_t **oPr)
{
    auto clear = [&]() {
        if (oCSize) *oCSize = 0;
        if (oRSize) *oRSize = 0;
        if (oFc) *oFc = nullptr;
        if (oFr) *oFr = nullptr;
        if (oBc) *oBc = nullptr;
        if (oBr) *oBr = nullptr;
        if (oPr) *oPr = nullptr;
    };
    clear();
    scc_host_graph g;
    std::memset(&g, 0, sizeof g);
    if (!filename || scc_load_gtgraph(filename, &g) != SCC_OK) {
        printf("Error Reading graph file\n");   // load.cpp:26
        return;
    }
    uint32_t *Pr = nullptr;
    const char *pfile = std::getenv("SCC_PARTITION_FILE");
    if (pfile && *pfile) {
        Pr = new uint32_t[(size_t)g.V + 2];
        if (scc_load_partition(pfile, g.V, Pr) != SCC_OK) {
            printf("Error Reading metis graph file\n");   // load.cpp:121
            delete[] Pr;
            scc_host_graph_free(&g);
            return;
        }
    }
    if (oCSize) *oCSize = g.E;
    if (oRSize) *oRSize = g.V;
    if (oFc) *oFc = copy_new(g.Fc, g.E);
    if (oFr) *oFr = copy_new(g.Fr, (size_t)g.V + 2);
    if (oBc) *oBc = copy_new(g.Bc, g.E);
    if (oBr) *oBr = copy_new(g.Br, (size_t)g.V + 2);
    if (oPr) *oPr = Pr;
    else delete[] Pr;
    scc_host_graph_free(&g);
}
