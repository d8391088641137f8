// mm_io.hip -- MatrixMarket files of the Newton system (host code): the
// reference's linear-system export (GlobalLinearSystem::
// export_matrix_market_files, linear_system/linear_system/
// global_linear_system.cu:98-261) and the readers of its solve-only mode
// (read_mm_coordinate_real, read_mm_array_vector_mx1,
// to_upper_triangle_block_triplets, StiffGIPC/main.cu:387-590).
//
// Formats (NIST MatrixMarket, written as the reference's mmio calls do):
//   A: "%%MatrixMarket matrix coordinate real general", "M N nnz", then
//      1-based "row col value" lines, every upper 3x3 block (row-major within
//      the block) followed by its transpose when off the diagonal;
//   b: "%%MatrixMarket matrix array real general", "M 1", one value a line.
// Values are printed with %.17g (round-trip exact for doubles).
#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "mm_io.h"
#include "sgipc.h"

namespace sgipc
{

namespace
{
struct File
{
    FILE* f = nullptr;
    explicit File(FILE* p)
        : f(p)
    {
    }
    ~File()
    {
        if(f)
            std::fclose(f);
    }
};

// mkdir -p
void make_dirs(const std::string& path)
{
    std::string cur;
    size_t      i = 0;
    while(i <= path.size())
    {
        size_t j = path.find('/', i);
        if(j == std::string::npos)
            j = path.size();
        cur = path.substr(0, j);
        if(!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            throw Error(SGIPC_EINVAL, "cannot create directory " + cur);
        i = j + 1;
    }
}

// next line that is not a comment (MatrixMarket comments start with '%')
bool data_line(FILE* f, std::string& line)
{
    char buf[512];
    while(std::fgets(buf, sizeof(buf), f))
    {
        const char* p = buf;
        while(*p == ' ' || *p == '\t')
            p++;
        if(*p == '%' || *p == '\n' || *p == '\r' || *p == 0)
            continue;
        line = buf;
        return true;
    }
    return false;
}

std::string lower(std::string s)
{
    for(char& c : s)
        c = (char)std::tolower((unsigned char)c);
    return s;
}

// banner words after "%%MatrixMarket": object, format, field, symmetry
std::vector<std::string> banner(FILE* f, const std::string& path)
{
    char buf[512];
    SG_CHECK(std::fgets(buf, sizeof(buf), f), SGIPC_EINVAL, "empty MatrixMarket file: " + path);
    std::vector<std::string> w;
    char                     t[5][128];
    int n = std::sscanf(buf, "%127s %127s %127s %127s %127s", t[0], t[1], t[2], t[3], t[4]);
    SG_CHECK(n == 5 && std::string(t[0]) == "%%MatrixMarket", SGIPC_EINVAL,
             "not a MatrixMarket banner: " + path);
    for(int k = 1; k < 5; k++)
        w.push_back(lower(t[k]));
    return w;
}
}  // namespace

void mm_write_system(const std::string& A_path, const std::string& b_path, int nb,
                     const std::vector<int>& rowptr, const std::vector<int>& cols,
                     const std::vector<double>& vals, const std::vector<double>& b)
{
    const int M = 3 * nb;
    SG_CHECK((int)b.size() == M, SGIPC_EINVAL, "b size mismatch with matrix rows");
    {
        File f(std::fopen(b_path.c_str(), "w"));
        SG_CHECK(f.f, SGIPC_EINVAL, "failed to open for write: " + b_path);
        std::fprintf(f.f, "%%%%MatrixMarket matrix array real general\n%d %d\n", M, 1);
        for(int i = 0; i < M; i++)
            std::fprintf(f.f, "%.17g\n", b[i]);
        SG_CHECK(!std::ferror(f.f), SGIPC_EINVAL, "failed to write " + b_path);
    }
    long long nnz = 0;
    for(int i = 0; i < nb; i++)
        for(int k = rowptr[i]; k < rowptr[i + 1]; k++)
            if(cols[k] >= i)
                nnz += cols[k] == i ? 9 : 18;
    SG_CHECK(nnz <= 0x7fffffffLL, SGIPC_ECAPACITY, "A.mtx too large (nnz exceeds int)");
    File f(std::fopen(A_path.c_str(), "w"));
    SG_CHECK(f.f, SGIPC_EINVAL, "failed to open for write: " + A_path);
    std::fprintf(f.f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n", M, M, (int)nnz);
    for(int bi = 0; bi < nb; bi++)
        for(int k = rowptr[bi]; k < rowptr[bi + 1]; k++)
        {
            const int bj = cols[k];
            if(bj < bi)
                continue;
            const double* B = vals.data() + 9 * (size_t)k;
            for(int i = 0; i < 3; i++)
                for(int j = 0; j < 3; j++)
                    std::fprintf(f.f, "%d %d %.17g\n", 3 * bi + i + 1, 3 * bj + j + 1, B[3 * i + j]);
            if(bi != bj)
                for(int i = 0; i < 3; i++)
                    for(int j = 0; j < 3; j++)
                        std::fprintf(f.f, "%d %d %.17g\n", 3 * bj + i + 1, 3 * bi + j + 1,
                                     B[3 * j + i]);
        }
    SG_CHECK(!std::ferror(f.f), SGIPC_EINVAL, "failed to write " + A_path);
}

std::string mm_system_paths(const std::string& dir, int frame, int newton_step, std::string* b_path)
{
    std::string base = dir + "/linear_system";
    make_dirs(base);
    char name[64];
    std::snprintf(name, sizeof(name), "frame_%05d_ns_%d", frame, newton_step);
    if(b_path)
        *b_path = base + "/" + name + "_b.mtx";
    return base + "/" + name + "_A.mtx";
}

void mm_read_upper_blocks(const std::string& path, int* nb, std::vector<int>& br,
                          std::vector<int>& bc, std::vector<double>& bv)
{
    File f(std::fopen(path.c_str(), "r"));
    SG_CHECK(f.f, SGIPC_EINVAL, "failed to open MatrixMarket file: " + path);
    auto w = banner(f.f, path);
    SG_CHECK(w[0] == "matrix" && w[1] == "coordinate" && (w[2] == "real" || w[2] == "double"),
             SGIPC_EINVAL, "A must be a coordinate real MatrixMarket matrix: " + path);
    SG_CHECK(w[3] == "general" || w[3] == "symmetric", SGIPC_EINVAL,
             "A must be general or symmetric: " + path);
    const bool  sym = w[3] == "symmetric";
    std::string line;
    SG_CHECK(data_line(f.f, line), SGIPC_EINVAL, "missing MatrixMarket size line: " + path);
    long long M = 0, N = 0, nz = 0;
    SG_CHECK(std::sscanf(line.c_str(), "%lld %lld %lld", &M, &N, &nz) == 3, SGIPC_EINVAL,
             "bad MatrixMarket size line: " + path);
    SG_CHECK(M == N, SGIPC_EINVAL, "A must be square");
    SG_CHECK(M % 3 == 0 && M > 0, SGIPC_EINVAL, "A row count must be divisible by 3 (block size)");
    const int                                   n = (int)(M / 3);
    std::map<std::pair<int, int>, std::array<double, 9>> blocks;
    std::vector<char>                           diag(n, 0);
    for(long long e = 0; e < nz; e++)
    {
        SG_CHECK(data_line(f.f, line), SGIPC_EINVAL, "MatrixMarket file ends early: " + path);
        long long r = 0, c = 0;
        double    v = 0;
        SG_CHECK(std::sscanf(line.c_str(), "%lld %lld %lg", &r, &c, &v) == 3, SGIPC_EINVAL,
                 "bad MatrixMarket entry: " + path);
        r -= 1;
        c -= 1;
        SG_CHECK(r >= 0 && c >= 0 && r < M && c < N, SGIPC_EINVAL,
                 "A contains out-of-range indices");
        int bi = (int)(r / 3), bj = (int)(c / 3), li = (int)(r % 3), lj = (int)(c % 3);
        if(bi > bj)
        {
            if(!sym)
                continue;  // a general file lists both triangles: keep the upper one
            std::swap(bi, bj);
            std::swap(li, lj);
        }
        auto it = blocks.find({bi, bj});
        if(it == blocks.end())
            it = blocks.emplace(std::make_pair(bi, bj), std::array<double, 9>{}).first;
        it->second[3 * li + lj] += v;
        if(sym && bi == bj && li != lj)
            it->second[3 * lj + li] += v;
        if(bi == bj)
            diag[bi] = 1;
    }
    for(int i = 0; i < n; i++)
        SG_CHECK(diag[i], SGIPC_EINVAL,
                 "A is missing diagonal 3x3 block at block index " + std::to_string(i));
    *nb = n;
    br.clear();
    bc.clear();
    bv.clear();
    for(auto& kv : blocks)  // (row, col) order
    {
        br.push_back(kv.first.first);
        bc.push_back(kv.first.second);
        bv.insert(bv.end(), kv.second.begin(), kv.second.end());
    }
}

std::vector<double> mm_read_vector(const std::string& path)
{
    File f(std::fopen(path.c_str(), "r"));
    SG_CHECK(f.f, SGIPC_EINVAL, "failed to open MatrixMarket file: " + path);
    auto w = banner(f.f, path);
    SG_CHECK(w[0] == "matrix" && w[1] == "array" && (w[2] == "real" || w[2] == "double"),
             SGIPC_EINVAL, "b must be an array real MatrixMarket matrix: " + path);
    std::string line;
    SG_CHECK(data_line(f.f, line), SGIPC_EINVAL, "missing MatrixMarket size line: " + path);
    long long M = 0, N = 0;
    SG_CHECK(std::sscanf(line.c_str(), "%lld %lld", &M, &N) == 2, SGIPC_EINVAL,
             "bad MatrixMarket size line: " + path);
    SG_CHECK(N == 1, SGIPC_EINVAL, "b must be a MatrixMarket array of shape (M x 1): " + path);
    std::vector<double> b((size_t)M);
    for(long long i = 0; i < M; i++)
    {
        SG_CHECK(data_line(f.f, line), SGIPC_EINVAL,
                 "failed to read b entry " + std::to_string(i) + " from: " + path);
        SG_CHECK(std::sscanf(line.c_str(), "%lg", &b[i]) == 1, SGIPC_EINVAL,
                 "bad b entry " + std::to_string(i) + " in: " + path);
    }
    return b;
}

void upper_bcoo_to_full_bsr(int nb, int nnz, const int* rows, const int* cols, const double* vals,
                            std::vector<int>& rowptr, std::vector<int>& out_cols,
                            std::vector<double>& out_vals)
{
    std::vector<std::map<int, std::array<double, 9>>> R(nb);
    for(int k = 0; k < nnz; k++)
    {
        int i = rows[k], j = cols[k];
        SG_CHECK(i >= 0 && j >= 0 && i < nb && j < nb, SGIPC_EINVAL, "block index out of range");
        SG_CHECK(i <= j, SGIPC_EINVAL, "upper BCOO blocks must have row <= col");
        const double* B = vals + 9 * (size_t)k;
        auto&         a = R[i][j];
        for(int e = 0; e < 9; e++)
            a[e] += B[e];
        if(i != j)
        {
            auto& t = R[j][i];
            for(int r = 0; r < 3; r++)
                for(int c = 0; c < 3; c++)
                    t[3 * r + c] += B[3 * c + r];
        }
    }
    rowptr.assign(nb + 1, 0);
    out_cols.clear();
    out_vals.clear();
    for(int i = 0; i < nb; i++)
    {
        SG_CHECK(R[i].count(i), SGIPC_EINVAL,
                 "missing diagonal block at block index " + std::to_string(i));
        for(auto& kv : R[i])
        {
            out_cols.push_back(kv.first);
            out_vals.insert(out_vals.end(), kv.second.begin(), kv.second.end());
        }
        rowptr[i + 1] = (int)out_cols.size();
    }
}

}  // namespace sgipc
