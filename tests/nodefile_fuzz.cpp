// Host-only driver of the node-file reader (se_rbf_projects_amd/csrc/nodefile.cpp), built by
// tests/test_nodefile.py with AddressSanitizer and UndefinedBehaviorSanitizer: every truncation
// of a node file, then seeded damage of up to three bytes in its first 11000 bytes (superblock
// and object headers).  Each variant must be read or refused; a read past the file buffer or
// the output arrays, or any undefined behaviour, stops the run with the sanitizer's report.
#include <cstdio>
#include <string>
#include <vector>

#include "serbf.h"

namespace serbf {
void set_error(const std::string &) {}
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <node file> <scratch file>\n", argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<unsigned char> d;
    for (int c; (c = std::fgetc(f)) != EOF;) d.push_back((unsigned char)c);
    std::fclose(f);
    const size_t damaged = 3000, span = d.size() < 11000 ? d.size() : 11000;
    std::vector<double> a(5 * 80);
    long ok = 0, refused = 0;
    unsigned s = 12345;
    for (size_t trial = 0; trial < d.size() + damaged; trial++) {
        std::vector<unsigned char> e = d;
        if (trial < d.size()) {
            e.resize(trial);
        } else {
            for (int k = 0; k < 3; k++) {
                s = s * 1103515245u + 12345u;
                const size_t i = (s >> 8) % span;
                s = s * 1103515245u + 12345u;
                e[i] = (unsigned char)(s >> 16);
            }
        }
        FILE *g = std::fopen(argv[2], "wb");
        if (!g) return 2;
        std::fwrite(e.data(), 1, e.size(), g);
        std::fclose(g);
        int N = 80;
        const int rc = serbf_read_nodeset(argv[2], &N, &a[0], &a[80], &a[160], &a[240], &a[320]);
        if (rc == SERBF_OK && (N < 1 || N > 80)) {
            std::fprintf(stderr, "variant %zu: accepted with N = %d\n", trial, N);
            return 1;
        }
        if (rc != SERBF_OK && rc != SERBF_ERR_ARG) {
            std::fprintf(stderr, "variant %zu: unexpected status %d\n", trial, rc);
            return 1;
        }
        rc ? refused++ : ok++;
    }
    std::printf("read %ld, refused %ld\n", ok, refused);
    return 0;
}
