// A caller written against the reference's own declarations: the prototypes
// below are those of include/scc.h:49-53 and include/load.h:6, declared here
// directly (C++ linkage, void), not taken from this library's headers, so a
// successful link shows that libscc_amd.so exports the reference's mangled
// names.  Used by the drop-in tests:
//   dropin_caller load <file.gr>             loadFullGraph only (no GPU)
//   dropin_caller run <g|h|x|y|d> <file.gr> [W] [p q]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdint.h>

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
void loadFullGraph(const char *filename, uint32_t *oCSize, uint32_t *oRSize, uint32_t **oFc,
                   uint32_t **oFr, uint32_t **oBc, uint32_t **oBr, uint32_t **oPr);

// FNV-1a over an array, so the test can compare the loaded arrays with the
// oracle's layout without shipping them through stdout
static uint64_t fnv(const uint32_t *p, size_t n)
{
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    const bool run = std::strcmp(argv[1], "run") == 0;
    const char *file = run ? argv[3] : argv[2];
    uint32_t CSize = 7, RSize = 7;
    uint32_t *Fc = nullptr, *Fr = nullptr, *Bc = nullptr, *Br = nullptr, *Pr = nullptr;
    loadFullGraph(file, &CSize, &RSize, &Fc, &Fr, &Bc, &Br, &Pr);
    if (!run) {
        if (!Fr) {
            printf("empty V=%u E=%u\n", RSize, CSize);
            return 3;
        }
        printf("V=%u E=%u Fr=%016llx Fc=%016llx Br=%016llx Bc=%016llx Pr=%s\n", RSize, CSize,
               (unsigned long long)fnv(Fr, (size_t)RSize + 2), (unsigned long long)fnv(Fc, CSize),
               (unsigned long long)fnv(Br, (size_t)RSize + 2), (unsigned long long)fnv(Bc, CSize),
               Pr ? "set" : "null");
    } else {
        const char algo = argv[2][0];
        const int warp = argc > 4 ? std::atoi(argv[4]) : 1;
        const bool t1 = argc > 5 ? argv[5][0] != '0' : true;
        const bool t2 = argc > 6 ? argv[6][0] != '0' : true;
        try {
            switch (algo) {
            case 'g': vHong(CSize, RSize, Fc, Fr, Bc, Br, t1, t2); break;
            case 'h': vSlota(CSize, RSize, Fc, Fr, Bc, Br, t1, t2); break;
            case 'x': wHong(CSize, RSize, Fc, Fr, Bc, Br, t1, t2, warp); break;
            case 'y': wSlota(CSize, RSize, Fc, Fr, Bc, Br, t1, t2, warp); break;
            default: detectSCC(CSize, RSize, Fc, Fr, Bc, Br, Pr, t1, t2); break;
            }
        } catch (const char *e) {
            printf("caught: %s\n", e);
            return 1;
        }
        printf("\n");
    }
    delete[] Fr;
    delete[] Fc;
    delete[] Br;
    delete[] Bc;
    delete[] Pr;
    return 0;
}
