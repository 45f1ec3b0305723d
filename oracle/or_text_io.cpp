// Oracle for the reference's text configuration files (test infrastructure
// only): the writer and reader restated with the reference's own mechanism,
// C++ iostreams, so that the library's threaded to_chars/from_chars
// implementation is checked against an independent formatter and parser.
//
//   or_text_io write nc Nx Ny Nz digits links|efield in.bin out.txt
//       in.bin: the host AoS array (doubles); out.txt as IO::SaveConfiguration
//       writes it (IO/SaveConfiguration.cpp:15-51): loops z, y, x, mu; a link
//       line is "x y z mu " followed by every value and a space
//       (MatrixToString, SUNc/SU2/GroupOperations.cpp:260-272,
//       SUNc/SU3/GroupOperations.cpp:539-551), a field line "x y z mu" and
//       " value" per component, both with the stream's precision set to digits.
//   or_text_io read nc Nx Ny Nz links|efield in.txt out.bin
//       as IO::LoadConfiguration reads it (IO/LoadConfiguration.cpp:46-72 for
//       links: getline, coordinates, then the values by operator>>; :84-116
//       for fields: x y z mu and the components by operator>>), out.bin the
//       host AoS array.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

static int fail(const char* m) {
    fprintf(stderr, "or_text_io: %s\n", m);
    return 2;
}

int main(int argc, char** argv) {
    if (argc < 8) return fail("usage: write nc Nx Ny Nz digits kind in out | read nc Nx Ny Nz kind in out");
    const std::string mode = argv[1];
    const int nc = atoi(argv[2]);
    const int N[3] = {atoi(argv[3]), atoi(argv[4]), atoi(argv[5])};
    const long V = (long)N[0]*N[1]*N[2];
    const bool write = mode == "write";
    if (write && argc != 10) return fail("write needs 9 arguments");
    if (!write && argc != 9) return fail("read needs 8 arguments");
    const int digits = write ? atoi(argv[6]) : 0;
    const std::string kind = argv[write ? 7 : 6];
    const char* in = argv[write ? 8 : 7];
    const char* out = argv[write ? 9 : 8];
    const bool links = kind == "links";
    const int nr = links ? (nc == 2 ? 4 : 18) : (nc == 2 ? 3 : 8);
    std::vector<double> A((size_t)(3*V*nr), 0.0);
    auto at = [&](int x, int y, int z, int mu) { return &A[(size_t)nr*(mu + 3*(x + N[0]*(y + (long)N[1]*z)))]; };
    if (write) {
        FILE* f = fopen(in, "rb");
        if (!f || fread(A.data(), sizeof(double), A.size(), f) != A.size()) return fail("cannot read the array");
        fclose(f);
        std::ofstream os(out);
        os.precision(digits);
        for (int z = 0; z < N[2]; z++)
            for (int y = 0; y < N[1]; y++)
                for (int x = 0; x < N[0]; x++)
                    for (int mu = 0; mu < 3; mu++) {
                        const double* v = at(x, y, z, mu);
                        if (links) {
                            std::stringstream m;
                            m.precision(digits);
                            for (int k = 0; k < nr; k++) m << v[k] << " ";
                            os << x << " " << y << " " << z << " " << mu << " " << m.str() << std::endl;
                        } else {
                            os << x << " " << y << " " << z << " " << mu;
                            for (int k = 0; k < nr; k++) os << " " << v[k];
                            os << std::endl;
                        }
                    }
        return os.good() ? 0 : fail("write failed");
    }
    std::ifstream is(in);
    if (!is) return fail("cannot open the text file");
    if (links) {
        std::string line;
        while (is.good()) {
            getline(is, line);
            if (line.empty()) continue;
            std::stringstream ls(line);
            int x, y, z, mu;
            ls >> x >> y >> z >> mu;
            double* v = at(x, y, z, mu);
            for (int k = 0; k < nr; k++) ls >> v[k];
        }
    } else {
        long count = 0;
        int x, y, z, mu;
        while (is >> x >> y >> z >> mu) {
            double* v = at(x, y, z, mu);
            for (int k = 0; k < nr; k++) is >> v[k];
            count++;
        }
        if (count != 3*V) return fail("wrong number of field records");
    }
    FILE* f = fopen(out, "wb");
    if (!f || fwrite(A.data(), sizeof(double), A.size(), f) != A.size()) return fail("cannot write the array");
    fclose(f);
    return 0;
}
