// scc_cli.cpp -- the `scc` executable: a drop-in for the reference CLI
// (src/scc.cu:21-115) on top of libscc_amd.so.
//
//   scc -a [g/h/x/y/d] -p [0/1] -q [0/1] -w [1/2/4/8/16/32/64] -f <file>
//
// Same flags, parsing and stdout lines as the reference.  Extra options
// (not in the reference): -P <partition file> for -a d (the reference reads
// a hard-coded path, load.cpp:118), -o <file> writes the canonical label of
// every vertex, one per line, -c <file> writes the condensation DAG (one
// node per SCC) as GTgraph text, -e <0..4> selects the frontier expansion,
// -D <device> selects the HIP device.  -f also takes a binary CSR cache.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>
#include <vector>

#include "scc_amd.h"

static void print_help()
{
    printf("To execute:\n");
    printf("./scc -a [g/h/x/y/d] -p [0/1] -q [0/1] -w [1/2/4/8/16/32/64] -f <file>\n");
    printf("-a algorith to use   g - vHong, h - vSlota, x - wHong, y - wSlota, d - detectSCC\n");
    printf("-p Trim-1 enable\n");
    printf("-q Trim-2 enable\n");
    printf("-w warp size\n");
    printf("-f input file\n");
    printf("Please note that the input graph must be in GTgraph format\n");
    printf("If not provided the results are undefined\n");
    printf("Extra options: -P <partition file> (-a d), -o <label file>, -c <condensation DAG file>,\n"
           "  -e <expand 0-4>, -D <device>; -f also reads a binary CSR cache\n");
}

int main(int argc, char **argv)
{
    if (argc < 11) {   // scc.cu:23-26
        print_help();
        return 1;
    }
    const char *file = nullptr, *part_file = nullptr, *label_file = nullptr, *dag_file = nullptr;
    char algo = 0;
    bool trim1 = true, trim2 = true;
    int warp = 1, expand = SCC_EXPAND_AUTO, device = 0;
    int c;
    while ((c = getopt(argc, argv, "a:p:q:w:f:P:o:c:e:D:")) != -1) {
        switch (c) {
        case 'a': algo = optarg[0]; break;
        case 'p': trim1 = optarg[0] != '0'; break;   // scc.cu:39-41
        case 'q': trim2 = optarg[0] != '0'; break;
        case 'w': warp = atoi(optarg); break;
        case 'f': file = optarg; break;
        case 'P': part_file = optarg; break;
        case 'o': label_file = optarg; break;
        case 'c': dag_file = optarg; break;
        case 'e': expand = atoi(optarg); break;
        case 'D': device = atoi(optarg); break;
        default: print_help(); return 1;
        }
    }
    if (algo != 'g' && algo != 'h' && algo != 'x' && algo != 'y' && algo != 'd') {
        print_help();   // scc.cu:99-101
        return 1;
    }
    if (!file) {
        printf("Error Reading graph file\n");
        return 1;
    }
    scc_host_graph g;
    std::memset(&g, 0, sizeof g);
    scc_status s = scc_load_graph(file, &g);
    if (s != SCC_OK) {
        printf("Error Reading graph file\n");
        fprintf(stderr, "%s\n", scc_last_error());
        return 1;
    }
    std::vector<uint32_t> part;
    if (algo == 'd') {
        if (!part_file) {
            printf("Error Reading metis graph file\n");   // load.cpp:121
            fprintf(stderr, "-a d needs -P <partition file>\n");
            scc_host_graph_free(&g);
            return 1;
        }
        part.resize((size_t)g.V + 2);
        if (scc_load_partition(part_file, g.V, part.data()) != SCC_OK) {
            printf("Error Reading metis graph file\n");
            fprintf(stderr, "%s\n", scc_last_error());
            scc_host_graph_free(&g);
            return 1;
        }
    }
    int rc;
    if (expand == SCC_EXPAND_AUTO && !label_file && !dag_file && device == 0) {
        switch (algo) {
        case 'g': s = scc_vHong(g.E, g.V, g.Fc, g.Fr, g.Bc, g.Br, trim1, trim2); break;
        case 'h': s = scc_vSlota(g.E, g.V, g.Fc, g.Fr, g.Bc, g.Br, trim1, trim2); break;
        case 'x': s = scc_wHong(g.E, g.V, g.Fc, g.Fr, g.Bc, g.Br, trim1, trim2, warp); break;
        case 'y': s = scc_wSlota(g.E, g.V, g.Fc, g.Fr, g.Bc, g.Br, trim1, trim2, warp); break;
        default: s = scc_detectSCC(g.E, g.V, g.Fc, g.Fr, g.Bc, g.Br, part.data(), trim1, trim2); break;
        }
        if (s != SCC_OK) printf("%s\n", scc_last_error());   // scc.cu:104-108
        rc = s == SCC_OK ? 0 : 1;
    } else {
        scc_ctx *ctx = nullptr;
        scc_opts o;
        scc_opts_default(&o);
        o.algo = algo;
        o.trim1 = trim1;
        o.trim2 = trim2;
        o.warp = warp;
        o.expand = expand;
        std::vector<uint32_t> lab((size_t)g.V + 1);
        scc_stats st;
        s = scc_ctx_create(device, &ctx);
        if (s == SCC_OK)
            s = scc_solve(ctx, &g, algo == 'd' ? part.data() : nullptr, &o, lab.data(), &st);
        if (s != SCC_OK) {
            printf("%s\n", scc_last_error());
            scc_ctx_destroy(ctx);
            scc_host_graph_free(&g);
            return 1;
        }
        if (algo == 'g') printf("numberOf1Sccs: %u\n", st.num_trim1);
        printf("\nNumber Of Sccs : %u", st.num_sccs);
        printf("\nTime : %f", st.ms_scc);
        if (algo == 'd') printf("\n");
        rc = 0;
        if (label_file) {
            FILE *f = fopen(label_file, "w");
            if (!f) {
                fprintf(stderr, "cannot write %s\n", label_file);
                rc = 1;
            } else {
                for (uint32_t v = 1; v <= g.V; v++) fprintf(f, "%u\n", lab[v]);
                fclose(f);
            }
        }
        if (dag_file) {
            scc_host_graph dag;
            std::memset(&dag, 0, sizeof dag);
            if (scc_condensation(ctx, &dag, nullptr) != SCC_OK || scc_write_gtgraph(dag_file, &dag) != SCC_OK) {
                fprintf(stderr, "%s\n", scc_last_error());
                rc = 1;
            }
            scc_host_graph_free(&dag);
        }
        scc_ctx_destroy(ctx);
    }
    scc_host_graph_free(&g);
    if (rc != 0) return 1;   // scc.cu:104-108
    printf("\n");
    return 0;
}
