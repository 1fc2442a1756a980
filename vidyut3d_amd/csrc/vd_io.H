// Plotfile and checkpoint I/O on the host (Source/Plot.cpp:32-225).
//
// Plotfile: the directory amrex::WriteMultiLevelPlotfile writes (AMReX's
// plotfile format "HyperCLaw-V1.1"):
//   <dir>/Header                 text: format name, variable count and names,
//                                dimension, time, finest level, problem lo
//                                and hi, refinement ratios, level domains,
//                                level steps, cell sizes per level,
//                                coordinate system, boundary width 0, then
//                                per level "lev ngrids time", its step, its
//                                boxes in physical coordinates (lo hi per
//                                direction) and the path of its MultiFab
//   <dir>/Level_<l>/Cell_H       VisMF header (version 1): version, layout
//                                (1 = NFiles), component count, ghost width,
//                                the BoxArray, the FabOnDisk entries (data
//                                file and byte offset per box), per-box
//                                minima and maxima of every component
//   <dir>/Level_<l>/Cell_D_00000 per box a text line
//                                "FAB <real descriptor><box> <ncomp>" and the
//                                box's values, x fastest, component
//                                outermost; the descriptor names IEEE double
//                                (64 11 52 0 1 12 0 1023) and the byte order,
//                                (8 7 6 5 4 3 2 1) for little-endian
// Checkpoint: the reference's own Header (WriteCheckpointFile: title,
// finest level, istep, dt and t_new per level, each level's BoxArray) and one
// VisMF MultiFab "phi" per level (<dir>/Level_<l>/phi_H, phi_D_00000).
//
// Writers produce one box per level (the context's layout).  Readers take
// any number of boxes and data files per level, ghost cells included, in
// either byte order, as AMReX writes them for a decomposed domain, and
// assemble the boxes into the level's one box.
#ifndef VD_IO_H_
#define VD_IO_H_

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

namespace vd {
namespace io {

struct IoError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

// A box in a level's own index space (inclusive bounds).
struct IBox
{
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    long long cells() const
    {
        return (long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
    }
};

// amrex::Concatenate: root followed by num with at least mindigits digits.
inline std::string concatenate(const std::string& root, int num, int mindigits = 5)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "%0*d", mindigits, num);
    return root + buf;
}

inline bool is_dir(const std::string& p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

inline void make_dirs(const std::string& p)
{
    size_t i = 0;
    while (i <= p.size())
    {
        size_t j = p.find('/', i);
        if (j == std::string::npos) j = p.size();
        const std::string cur = p.substr(0, j);
        if (!cur.empty() && !is_dir(cur) && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            throw IoError("cannot create directory " + cur + ": " + std::strerror(errno));
        i = j + 1;
    }
}

// The output directory with one subdirectory per level.  An existing
// directory of that name is first renamed to <dir>.old.<pid>.<n>, as AMReX
// renames an existing plotfile or checkpoint, so nothing is overwritten.
inline void prepare_dir(const std::string& dir, int nlevels)
{
    if (dir.empty()) throw IoError("empty output directory name");
    if (is_dir(dir))
    {
        static int counter = 0;
        const std::string old = dir + ".old." + std::to_string((long)getpid()) + "." + std::to_string(counter++);
        if (std::rename(dir.c_str(), old.c_str()) != 0)
            throw IoError("cannot rename the existing " + dir + ": " + std::strerror(errno));
    }
    make_dirs(dir);
    for (int l = 0; l < nlevels; ++l) make_dirs(dir + "/Level_" + std::to_string(l));
}

inline std::string ivec(const int v[3], int dim)
{
    std::string s = "(";
    for (int d = 0; d < dim; ++d) s += (d ? "," : "") + std::to_string(v[d]);
    return s + ")";
}

// AMReX's Box text form: ((lo) (hi) (type)), cell-centred type 0.
inline std::string box_str(const IBox& b, int dim)
{
    const int z[3] = {0, 0, 0};
    return "(" + ivec(b.lo, dim) + " " + ivec(b.hi, dim) + " " + ivec(z, dim) + ")";
}

// BoxArray::writeOn: "(n 0", the boxes one per line, ")".
inline void write_boxarray(std::ostream& os, const std::vector<IBox>& ba, int dim)
{
    os << '(' << ba.size() << ' ' << 0 << '\n';
    for (const IBox& b : ba) os << box_str(b, dim) << '\n';
    os << ')';
}

// Integers of a text fragment, every non-digit (other than '-') a separator.
inline std::vector<long long> ints_of(const std::string& s)
{
    std::vector<long long> v;
    size_t i = 0;
    while (i < s.size())
    {
        if (std::isdigit((unsigned char)s[i]) || (s[i] == '-' && i + 1 < s.size() && std::isdigit((unsigned char)s[i + 1])))
        {
            size_t j = i + 1;
            while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j;
            v.push_back(std::stoll(s.substr(i, j - i)));
            i = j;
        } else
            ++i;
    }
    return v;
}

inline IBox parse_box(const std::string& s, int dim)
{
    const std::vector<long long> v = ints_of(s);
    if ((int)v.size() < 2 * dim) throw IoError("malformed box: " + s);
    IBox b;
    for (int d = 0; d < dim; ++d)
    {
        b.lo[d] = (int)v[d];
        b.hi[d] = (int)v[dim + d];
    }
    return b;
}

// Reads a BoxArray written by write_boxarray (or BoxArray::writeOn).
inline std::vector<IBox> read_boxarray(std::istream& is, int dim)
{
    std::string line;
    while (std::getline(is, line) && line.find('(') == std::string::npos)
    {
    }
    const std::vector<long long> head = ints_of(line);
    if (head.empty()) throw IoError("malformed BoxArray");
    std::vector<IBox> ba((size_t)head[0]);
    for (auto& b : ba)
    {
        if (!std::getline(is, line)) throw IoError("truncated BoxArray");
        b = parse_box(line, dim);
    }
    std::getline(is, line);   // ")"
    return ba;
}

inline const char* kRealDescLittle = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))";

inline bool host_little_endian()
{
    const uint16_t one = 1;
    unsigned char c;
    std::memcpy(&c, &one, 1);
    return c == 1;
}

inline std::string dir_of(const std::string& p)
{
    const size_t k = p.rfind('/');
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}
inline std::string base_of(const std::string& p)
{
    const size_t k = p.rfind('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

inline void write_text(const std::string& path, const std::string& text)
{
    std::ofstream f(path, std::ios::out | std::ios::trunc | std::ios::binary);
    if (!f) throw IoError("cannot open " + path + " for writing");
    f << text;
    f.close();
    if (!f) throw IoError("write error on " + path);
}

// VisMF::Write of a MultiFab of one box without ghosts: <prefix>_H and
// <prefix>_D_00000.  data: ncomp * cells values, x fastest, component
// outermost.
inline void write_vismf(const std::string& prefix, int ncomp, int dim, const IBox& b, const double* data)
{
    if (!host_little_endian()) throw IoError("big-endian hosts are not supported by the writer");
    const long long n = b.cells();
    const std::string dname = prefix + "_D_00000";
    {
        std::ofstream d(dname, std::ios::out | std::ios::trunc | std::ios::binary);
        if (!d) throw IoError("cannot open " + dname + " for writing");
        d << "FAB " << kRealDescLittle << box_str(b, dim) << ' ' << ncomp << '\n';
        d.write(reinterpret_cast<const char*>(data), (std::streamsize)(sizeof(double) * n * ncomp));
        d.close();
        if (!d) throw IoError("write error on " + dname);
    }
    std::vector<double> mn(ncomp), mx(ncomp);
    for (int c = 0; c < ncomp; ++c)
    {
        double lo = std::numeric_limits<double>::infinity(), hi = -lo;
        const double* v = data + (long long)c * n;
        for (long long q = 0; q < n; ++q)
        {
            lo = std::min(lo, v[q]);
            hi = std::max(hi, v[q]);
        }
        mn[c] = lo;
        mx[c] = hi;
    }
    std::ostringstream h;
    h << 1 << '\n' << 1 << '\n' << ncomp << '\n' << 0 << '\n';
    write_boxarray(h, {b}, dim);
    h << '\n';
    h << 1 << '\n' << "FabOnDisk: " << base_of(dname) << ' ' << 0 << '\n' << '\n';
    h.setf(std::ios::scientific, std::ios::floatfield);
    h.precision(17);
    h << 1 << ',' << ncomp << '\n';
    for (int c = 0; c < ncomp; ++c) h << mn[c] << ',';
    h << '\n';
    h << 1 << ',' << ncomp << '\n';
    for (int c = 0; c < ncomp; ++c) h << mx[c] << ',';
    h << '\n';
    write_text(prefix + "_H", h.str());
}

// VisMF::Read into one level box: every box of the MultiFab (grown by its
// ghost width in the data file) is read and its valid cells inside 'level'
// copied to out (ncomp * level.cells(), x fastest, component outermost);
// every cell of 'level' must be covered.
inline void read_vismf(const std::string& prefix, int ncomp, int dim, const IBox& level, double* out)
{
    std::ifstream h(prefix + "_H");
    if (!h) throw IoError("cannot open " + prefix + "_H");
    int vers = 0, how = 0, nc = 0;
    std::string ngline;
    h >> vers >> how >> nc;
    h >> std::ws;
    std::getline(h, ngline);
    if (vers != 1 || nc <= 0) throw IoError(prefix + "_H: unsupported VisMF header (version " + std::to_string(vers) + ")");
    if (nc != ncomp)
        throw IoError(prefix + "_H holds " + std::to_string(nc) + " components, the context " + std::to_string(ncomp));
    const std::vector<long long> ngv = ints_of(ngline);
    int ng[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) ng[d] = ngv.empty() ? 0 : (int)ngv[std::min<size_t>(d, ngv.size() - 1)];
    const std::vector<IBox> ba = read_boxarray(h, dim);
    size_t nfod = 0;
    h >> nfod;
    if (nfod != ba.size()) throw IoError(prefix + "_H: FabOnDisk count differs from the box count");
    std::vector<std::pair<std::string, long long>> fod(nfod);
    for (auto& f : fod)
    {
        std::string tag;
        h >> tag >> f.first >> f.second;
        if (tag != "FabOnDisk:") throw IoError(prefix + "_H: malformed FabOnDisk entry");
    }
    const long long nl = level.cells();
    std::vector<char> covered((size_t)nl, 0);
    const int nxl = level.hi[0] - level.lo[0] + 1, nyl = level.hi[1] - level.lo[1] + 1;
    const bool little = host_little_endian();
    for (size_t i = 0; i < ba.size(); ++i)
    {
        const std::string dpath = dir_of(prefix) + "/" + fod[i].first;
        std::ifstream d(dpath, std::ios::in | std::ios::binary);
        if (!d) throw IoError("cannot open " + dpath);
        d.seekg(fod[i].second);
        std::string fh;
        std::getline(d, fh);
        if (fh.compare(0, 4, "FAB ") != 0) throw IoError(dpath + ": no FAB header at offset " + std::to_string(fod[i].second));
        // "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (o1 .. o8)))((lo) (hi) (t)) ncomp"
        const size_t e = fh.find(")))");
        if (e == std::string::npos) throw IoError(dpath + ": malformed FAB header");
        const std::vector<long long> desc = ints_of(fh.substr(4, e + 3 - 4));
        if (desc.size() != 18 || desc[0] != 8 || desc[1] != 64 || desc[9] != 8)
            throw IoError(dpath + ": only 8-byte IEEE doubles are read (" + fh.substr(0, e + 3) + ")");
        bool file_little = true, file_big = true;
        for (int k = 0; k < 8; ++k)
        {
            file_little &= desc[10 + k] == 8 - k;
            file_big &= desc[10 + k] == k + 1;
        }
        if (!file_little && !file_big) throw IoError(dpath + ": unsupported byte order");
        const std::string rest = fh.substr(e + 3);
        const size_t bend = rest.rfind(')');
        const IBox fb = parse_box(rest.substr(0, bend + 1), dim);
        const std::vector<long long> tail = ints_of(rest.substr(bend + 1));
        if (tail.empty() || tail[0] != ncomp) throw IoError(dpath + ": FAB component count differs");
        const long long fn = fb.cells();
        std::vector<double> buf((size_t)(fn * ncomp));
        d.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(sizeof(double) * buf.size()));
        if (!d) throw IoError(dpath + ": truncated FAB data");
        if (file_little != little)
            for (double& v : buf)
            {
                unsigned char* p = reinterpret_cast<unsigned char*>(&v);
                std::reverse(p, p + 8);
            }
        // valid box = the BoxArray entry; the FAB may be grown by ng
        const IBox& vb = ba[i];
        for (int d2 = 0; d2 < dim; ++d2)
            if (fb.lo[d2] != vb.lo[d2] - ng[d2] || fb.hi[d2] != vb.hi[d2] + ng[d2])
                throw IoError(dpath + ": FAB box does not match the BoxArray");
        const int fnx = fb.hi[0] - fb.lo[0] + 1, fny = fb.hi[1] - fb.lo[1] + 1;
        for (int k = std::max(vb.lo[2], level.lo[2]); k <= std::min(vb.hi[2], level.hi[2]); ++k)
            for (int j = std::max(vb.lo[1], level.lo[1]); j <= std::min(vb.hi[1], level.hi[1]); ++j)
                for (int ii = std::max(vb.lo[0], level.lo[0]); ii <= std::min(vb.hi[0], level.hi[0]); ++ii)
                {
                    const long long src = (long long)(ii - fb.lo[0]) +
                                          (long long)fnx * ((long long)(j - fb.lo[1]) + (long long)fny * (k - fb.lo[2]));
                    const long long dst = (long long)(ii - level.lo[0]) +
                                          (long long)nxl * ((long long)(j - level.lo[1]) + (long long)nyl * (k - level.lo[2]));
                    for (int c = 0; c < ncomp; ++c) out[(long long)c * nl + dst] = buf[(size_t)((long long)c * fn + src)];
                    covered[(size_t)dst] = 1;
                }
    }
    if (std::find(covered.begin(), covered.end(), 0) != covered.end())
        throw IoError(prefix + ": the boxes do not cover the level's box " + box_str(level, dim));
}

// One level of a plotfile or checkpoint.
struct Level
{
    IBox box;           // its one box (level index space)
    IBox domain;        // the level's domain
    double dx[3];
    const double* data; // ncomp * box.cells() (writers)
};

// amrex::WriteMultiLevelPlotfile with refinement ratio 2 between levels.
inline void write_plotfile(const std::string& dir, const std::vector<std::string>& names, int dim, double time,
                           const double problo[3], const double probhi[3], const std::vector<Level>& lv,
                           const std::vector<int>& level_steps, int coord)
{
    const int nl = (int)lv.size();
    prepare_dir(dir, nl);
    std::ostringstream h;
    h.precision(17);
    h << "HyperCLaw-V1.1" << '\n' << names.size() << '\n';
    for (const auto& s : names) h << s << '\n';
    h << dim << '\n' << time << '\n' << nl - 1 << '\n';
    for (int d = 0; d < dim; ++d) h << problo[d] << ' ';
    h << '\n';
    for (int d = 0; d < dim; ++d) h << probhi[d] << ' ';
    h << '\n';
    for (int l = 0; l < nl - 1; ++l) h << 2 << ' ';
    h << '\n';
    for (int l = 0; l < nl; ++l) h << box_str(lv[l].domain, dim) << ' ';
    h << '\n';
    for (int l = 0; l < nl; ++l) h << level_steps[l] << ' ';
    h << '\n';
    for (int l = 0; l < nl; ++l)
    {
        for (int d = 0; d < dim; ++d) h << lv[l].dx[d] << ' ';
        h << '\n';
    }
    h << coord << '\n' << "0\n";
    for (int l = 0; l < nl; ++l)
    {
        h << l << ' ' << 1 << ' ' << time << '\n' << level_steps[l] << '\n';
        for (int d = 0; d < dim; ++d)
            h << problo[d] + (lv[l].box.lo[d] - lv[l].domain.lo[d]) * lv[l].dx[d] << ' '
              << problo[d] + (lv[l].box.hi[d] + 1 - lv[l].domain.lo[d]) * lv[l].dx[d] << '\n';
        h << "Level_" << l << "/Cell" << '\n';
    }
    write_text(dir + "/Header", h.str());
    for (int l = 0; l < nl; ++l)
        write_vismf(dir + "/Level_" + std::to_string(l) + "/Cell", (int)names.size(), dim, lv[l].box, lv[l].data);
}

// Vidyut::WriteCheckpointFile (Source/Plot.cpp:46-135).
inline void write_checkpoint(const std::string& dir, int ncomp, int dim, const std::vector<Level>& lv,
                             const std::vector<int>& istep, const std::vector<double>& dt,
                             const std::vector<double>& t_new)
{
    const int nl = (int)lv.size();
    prepare_dir(dir, nl);
    std::ostringstream h;
    h.precision(17);
    h << "Checkpoint file for Vidyut\n" << nl - 1 << "\n";
    for (int v : istep) h << v << " ";
    h << "\n";
    for (double v : dt) h << v << " ";
    h << "\n";
    for (double v : t_new) h << v << " ";
    h << "\n";
    for (int l = 0; l < nl; ++l)
    {
        write_boxarray(h, {lv[l].box}, dim);
        h << '\n';
    }
    write_text(dir + "/Header", h.str());
    for (int l = 0; l < nl; ++l) write_vismf(dir + "/Level_" + std::to_string(l) + "/phi", ncomp, dim, lv[l].box, lv[l].data);
}

// Vidyut::ReadCheckpointFile's Header (Source/Plot.cpp:137-217).
struct CheckpointHeader
{
    int finest_level = 0;
    std::vector<int> istep;
    std::vector<double> dt, t_new;
    std::vector<std::vector<IBox>> boxes;   // per level
};

inline CheckpointHeader read_checkpoint_header(const std::string& dir, int dim)
{
    std::ifstream is(dir + "/Header");
    if (!is) throw IoError("cannot open " + dir + "/Header");
    CheckpointHeader H;
    std::string line;
    std::getline(is, line);   // title
    if (!(is >> H.finest_level) || H.finest_level < 0) throw IoError(dir + "/Header: no finest level");
    std::getline(is, line);
    auto reals = [&](std::vector<double>& v) {
        std::getline(is, line);
        std::istringstream ls(line);
        std::string w;
        while (ls >> w) v.push_back(std::stod(w));
    };
    {
        std::vector<double> s;
        reals(s);
        for (double x : s) H.istep.push_back((int)x);
    }
    reals(H.dt);
    reals(H.t_new);
    if ((int)H.istep.size() <= H.finest_level || (int)H.t_new.size() <= H.finest_level)
        throw IoError(dir + "/Header: istep or t_new shorter than the level count");
    for (int l = 0; l <= H.finest_level; ++l) H.boxes.push_back(read_boxarray(is, dim));
    return H;
}

} // namespace io
} // namespace vd
#endif
