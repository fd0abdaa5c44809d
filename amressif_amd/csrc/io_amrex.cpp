// Checkpoint, plotfile and initial-state entry points (host C++).
//
// The reference writes its restart state with AMReX's VisMF
// (SaveCheckpoint / LoadCheckpoint, Source/src1_rever/utilities.cpp:792-1022)
// and its flow fields with WriteSingleLevelPlotfile (Export_Flow_Field,
// utilities.cpp:534-570).  AMReX is not part of this build, so this file
// writes and reads those on-disk layouts itself:
//
//   checkpointNNNNN/Header              title line, time, BoxArray
//   checkpointNNNNN/Level_0/<mf>_H      VisMF header: version, how, ncomp,
//                                       ngrow, BoxArray, FabOnDisk list, min/max
//   checkpointNNNNN/Level_0/<mf>_D_00000  per box: "FAB <real descriptor><box> ncomp\n"
//                                       then the box's doubles (i fastest, comp last)
//   pltNNNNN/Header                     HyperCLaw-V1.1 plotfile header
//   pltNNNNN/Level_0/Cell_H, Cell_D_00000
//
// The BoxArray is the domain chopped at the inputs' max_grid_size
// (main.cpp:317-320).  The reader takes any BoxArray whose boxes tile the
// domain and any FAB data file offsets, in IEEE double or float of either
// byte order, as AMReX's own reader does (ParallelCopy onto the solver's
// boxes, utilities.cpp:1007-1015).
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace {

struct Box2 {
    int lo[2], hi[2], typ[2];
};

// BoxArray::maxSize: a side of length L > m is cut into k = ceil(L / m)
// pieces as even as possible, the first L mod k one cell longer; x fastest.
std::vector<Box2> max_size_boxes(int n, int mx, int my)
{
    auto cuts = [](int len, int m) {
        std::vector<std::pair<int, int>> out;
        const int k = (len + m - 1) / m, size = len / k, extra = len % k;
        int lo = 0;
        for (int b = 0; b < k; ++b) {
            const int w = size + (b < extra ? 1 : 0);
            out.push_back({lo, lo + w - 1});
            lo += w;
        }
        return out;
    };
    std::vector<Box2> boxes;
    for (auto y : cuts(n, my > 0 ? my : n))
        for (auto x : cuts(n, mx > 0 ? mx : n)) boxes.push_back(Box2{{x.first, y.first}, {x.second, y.second}, {0, 0}});
    return boxes;
}

std::string box_str(const Box2 &b)
{
    std::ostringstream os;
    os << "((" << b.lo[0] << ',' << b.lo[1] << ") (" << b.hi[0] << ',' << b.hi[1] << ") (" << b.typ[0] << ','
       << b.typ[1] << "))";
    return os.str();
}

void write_boxarray(std::ostream &os, const std::vector<Box2> &boxes)
{
    os << '(' << boxes.size() << ' ' << 0 << '\n';
    for (const Box2 &b : boxes) os << box_str(b) << '\n';
    os << ')';
}

// ---- a minimal tokenizer over the text parts of the AMReX formats
struct Scan {
    const std::string &s;
    size_t k = 0;
    bool ok = true;
    explicit Scan(const std::string &str) : s(str) {}
    void ws()
    {
        while (k < s.size() && std::isspace((unsigned char)s[k])) ++k;
    }
    bool lit(char ch)
    {
        ws();
        if (k < s.size() && s[k] == ch) {
            ++k;
            return true;
        }
        ok = false;
        return false;
    }
    long integer()
    {
        ws();
        char *end = nullptr;
        const long v = std::strtol(s.c_str() + k, &end, 10);
        if (end == s.c_str() + k) ok = false;
        k = end - s.c_str();
        return v;
    }
    double real()
    {
        ws();
        char *end = nullptr;
        const double v = std::strtod(s.c_str() + k, &end);
        if (end == s.c_str() + k) ok = false;
        k = end - s.c_str();
        return v;
    }
    std::string word()
    {
        ws();
        const size_t b = k;
        while (k < s.size() && !std::isspace((unsigned char)s[k])) ++k;
        return s.substr(b, k - b);
    }
    void line()
    {
        while (k < s.size() && s[k] != '\n') ++k;
        if (k < s.size()) ++k;
    }
    // "(a,b)" in 2D; a third entry (3D files) is not accepted
    void ivect(int v[2])
    {
        lit('(');
        v[0] = (int)integer();
        lit(',');
        v[1] = (int)integer();
        lit(')');
    }
    Box2 box()
    {
        Box2 b{};
        lit('(');
        ivect(b.lo);
        ivect(b.hi);
        ivect(b.typ);
        lit(')');
        return b;
    }
    std::vector<Box2> boxarray()
    {
        std::vector<Box2> out;
        lit('(');
        const long nb = integer();
        integer();  // hash, written as 0
        for (long i = 0; ok && i < nb; ++i) out.push_back(box());
        lit(')');
        return out;
    }
};

bool read_file(const std::string &path, std::string &out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

bool make_dir(const std::string &p)
{
    if (mkdir(p.c_str(), 0755) == 0) return true;
    struct stat st;
    return errno == EEXIST && stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

// One global field of the one-box layout: rows j, row length w; (i, j) of
// the field's own index space (faces nodal in their direction).
struct Field {
    const double *a;  // writer
    double *out;      // reader
    int w, rows;
    int nodal;  // -1 cell, 0 x-nodal, 1 y-nodal
};

const char *kDoubleLE = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))";

// VisMF::Write of one MultiFab of ncomp fields over `boxes` (cell boxes; the
// fields' nodality turns them into face boxes), ngrow 0, NFiles form with
// one data file.  Returns false on an I/O failure.
bool write_vismf(const std::string &prefix, const std::vector<Box2> &cboxes, const std::vector<Field> &comps,
                 std::string &err)
{
    const int ncomp = (int)comps.size();
    const int nodal = comps[0].nodal;
    std::vector<Box2> boxes = cboxes;
    for (Box2 &b : boxes)
        if (nodal >= 0) {
            b.hi[nodal] += 1;
            b.typ[nodal] = 1;
        }
    const size_t slash = prefix.find_last_of('/');
    const std::string base = slash == std::string::npos ? prefix : prefix.substr(slash + 1);
    const std::string dname = prefix + "_D_00000";
    std::ofstream d(dname, std::ios::binary | std::ios::trunc);
    if (!d) {
        err = "cannot open " + dname;
        return false;
    }
    std::vector<long> offsets;
    std::vector<std::vector<double>> mins, maxs;
    std::vector<double> buf;
    for (const Box2 &b : boxes) {
        offsets.push_back((long)d.tellp());
        d << "FAB " << kDoubleLE << box_str(b) << ' ' << ncomp << '\n';
        const int wx = b.hi[0] - b.lo[0] + 1, wy = b.hi[1] - b.lo[1] + 1;
        buf.resize((size_t)wx * wy * ncomp);
        std::vector<double> mn(ncomp), mx(ncomp);
        for (int c = 0; c < ncomp; ++c) {
            double lo = HUGE_VAL, hi = -HUGE_VAL;
            for (int j = 0; j < wy; ++j)
                for (int i = 0; i < wx; ++i) {
                    const double x = comps[c].a[(size_t)(b.lo[0] + i) + (size_t)comps[c].w * (b.lo[1] + j)];
                    buf[(size_t)i + (size_t)wx * j + (size_t)wx * wy * c] = x;
                    lo = x < lo ? x : lo;
                    hi = x > hi ? x : hi;
                }
            mn[c] = lo;
            mx[c] = hi;
        }
        d.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)(buf.size() * sizeof(double)));
        mins.push_back(mn);
        maxs.push_back(mx);
    }
    d.close();
    if (!d) {
        err = "write failed: " + dname;
        return false;
    }
    std::ofstream h(prefix + "_H", std::ios::trunc);
    if (!h) {
        err = "cannot open " + prefix + "_H";
        return false;
    }
    h << 1 << '\n' << 1 << '\n' << ncomp << '\n' << 0 << '\n';  // Version_v1, NFiles, ncomp, ngrow
    write_boxarray(h, boxes);
    h << '\n' << boxes.size() << '\n';
    for (size_t b = 0; b < boxes.size(); ++b) h << "FabOnDisk: " << base << "_D_00000 " << offsets[b] << '\n';
    h.precision(17);
    for (const auto *mm : {&mins, &maxs}) {
        h << mm->size() << ',' << ncomp << '\n';
        for (const auto &row : *mm) {
            for (double x : row) h << x << ',';
            h << '\n';
        }
    }
    h.close();
    if (!h) {
        err = "write failed: " + prefix + "_H";
        return false;
    }
    return true;
}

// VisMF::Read of one MultiFab into global fields (each box's valid data is
// placed at its global position; a face shared by two boxes is taken from the
// later one).  comps[c].out receives component c.
bool read_vismf(const std::string &prefix, std::vector<Field> &comps, int n, std::string &err)
{
    std::string text;
    if (!read_file(prefix + "_H", text)) {
        err = "cannot read " + prefix + "_H";
        return false;
    }
    Scan sc(text);
    sc.integer();  // version
    sc.integer();  // how
    const long ncomp = sc.integer();
    // ngrow: an int or an IntVect
    sc.ws();
    int ngrow = 0;
    if (sc.k < text.size() && text[sc.k] == '(') {
        int v[2];
        sc.ivect(v);
        ngrow = v[0];
    } else {
        ngrow = (int)sc.integer();
    }
    std::vector<Box2> boxes = sc.boxarray();
    const long nfod = sc.integer();
    if (!sc.ok || ncomp < (long)comps.size() || nfod != (long)boxes.size()) {
        err = "malformed VisMF header " + prefix + "_H";
        return false;
    }
    const size_t slash = prefix.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : prefix.substr(0, slash);
    std::vector<std::string> files(nfod);
    std::vector<long> offs(nfod);
    for (long b = 0; b < nfod; ++b) {
        if (sc.word() != "FabOnDisk:") {
            err = "malformed FabOnDisk entry in " + prefix + "_H";
            return false;
        }
        files[b] = sc.word();
        offs[b] = sc.integer();
    }
    const int nodal = comps[0].nodal;
    std::string cur_name, data;
    for (long b = 0; b < nfod; ++b) {
        if (files[b] != cur_name) {
            if (!read_file(dir + "/" + files[b], data)) {
                err = "cannot read " + dir + "/" + files[b];
                return false;
            }
            cur_name = files[b];
        }
        if (offs[b] < 0 || (size_t)offs[b] >= data.size()) {
            err = "FAB offset outside " + files[b];
            return false;
        }
        const size_t eol = data.find('\n', (size_t)offs[b]);
        if (eol == std::string::npos) {
            err = "truncated FAB header in " + files[b];
            return false;
        }
        const std::string head = data.substr((size_t)offs[b], eol - (size_t)offs[b]);
        // "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))((lo) (hi) (typ)) ncomp"
        Scan fh(head);
        if (fh.word().rfind("FAB", 0) != 0) {
            err = "not a FAB header in " + files[b];
            return false;
        }
        fh.k = head.find('(');
        fh.lit('(');
        fh.lit('(');
        const long nbytes = fh.integer();
        fh.lit(',');
        fh.lit('(');
        for (int q = 0; q < 8; ++q) fh.integer();
        fh.lit(')');
        fh.lit(')');
        fh.lit(',');
        fh.lit('(');
        fh.integer();
        fh.lit(',');
        fh.lit('(');
        std::vector<int> order(nbytes > 0 && nbytes <= 8 ? nbytes : 0);
        for (auto &o : order) o = (int)fh.integer();
        fh.lit(')');
        fh.lit(')');
        fh.lit(')');
        const Box2 fb = fh.box();
        const long fcomp = fh.integer();
        if (!fh.ok || (nbytes != 8 && nbytes != 4) || fcomp < (long)comps.size()) {
            err = "unsupported FAB header in " + files[b] + ": " + head;
            return false;
        }
        // byte order: 1..nbytes is big-endian, nbytes..1 little-endian (this host)
        const bool swap = order.front() == 1;
        const int wx = fb.hi[0] - fb.lo[0] + 1, wy = fb.hi[1] - fb.lo[1] + 1;
        const size_t npts = (size_t)wx * wy;
        if (eol + 1 + npts * fcomp * nbytes > data.size()) {
            err = "truncated FAB data in " + files[b];
            return false;
        }
        const unsigned char *raw = reinterpret_cast<const unsigned char *>(data.data()) + eol + 1;
        // the valid region of this box (its grown cells, if any, are skipped)
        const int vlo0 = fb.lo[0] + ngrow, vlo1 = fb.lo[1] + ngrow, vhi0 = fb.hi[0] - ngrow, vhi1 = fb.hi[1] - ngrow;
        const int gw = nodal == 0 ? n + 1 : n, gh = nodal == 1 ? n + 1 : n;
        if (vlo0 < 0 || vlo1 < 0 || vhi0 >= gw || vhi1 >= gh || (nodal >= 0 && fb.typ[nodal] != 1)) {
            err = "box " + box_str(fb) + " of " + prefix + " does not fit the domain";
            return false;
        }
        for (size_t c = 0; c < comps.size(); ++c)
            for (int j = vlo1; j <= vhi1; ++j)
                for (int i = vlo0; i <= vhi0; ++i) {
                    const size_t e = (size_t)(i - fb.lo[0]) + (size_t)wx * (j - fb.lo[1]) + npts * c;
                    unsigned char bytes[8];
                    std::memcpy(bytes, raw + e * nbytes, nbytes);
                    if (swap)
                        for (int q = 0; q < nbytes / 2; ++q) std::swap(bytes[q], bytes[nbytes - 1 - q]);
                    double x;
                    if (nbytes == 8) {
                        std::memcpy(&x, bytes, 8);
                    } else {
                        float f;
                        std::memcpy(&f, bytes, 4);
                        x = f;
                    }
                    comps[c].out[(size_t)i + (size_t)comps[c].w * j] = x;
                }
    }
    return true;
}

std::string step_name(const char *prefix, int step)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "%05d", step);
    return std::string(prefix) + buf;
}

// WriteSingleLevelPlotfile of cell fields (n x n each, i fastest) named
// `names` into <dir>/<name>NNNNN: HyperCLaw-V1.1 Header and Level_0/Cell
int write_plot(amressif_ctx *ctx, const char *dir, const char *name, int step, double time,
               const std::vector<std::string> &names, const std::vector<const double *> &data)
{
    const amr::Ctx &c = ctx->c;
    const int n = c.g.n;
    const std::string root = std::string(dir) + "/" + step_name(name, step);
    if (!make_dir(root) || !make_dir(root + "/Level_0"))
        return amr_fail(ctx, AMRESSIF_ERR_ARG, "cannot create " + root + "/Level_0");
    const std::vector<Box2> boxes = max_size_boxes(n, c.in.max_grid_size[0], c.in.max_grid_size[1]);
    const double lo[2] = {c.in.lo_phy_dim[0], c.in.lo_phy_dim[1]}, hi[2] = {c.in.hi_phy_dim[0], c.in.hi_phy_dim[1]};
    const double dx[2] = {c.g.dx, c.g.dy};
    {
        std::ofstream h(root + "/Header", std::ios::trunc);
        if (!h) return amr_fail(ctx, AMRESSIF_ERR_ARG, "cannot open " + root + "/Header");
        h.precision(17);
        h << "HyperCLaw-V1.1\n" << names.size() << '\n';
        for (const std::string &nm : names) h << nm << '\n';
        h << 2 << '\n' << time << '\n' << 0 << '\n';
        h << lo[0] << ' ' << lo[1] << ' ' << '\n' << hi[0] << ' ' << hi[1] << ' ' << '\n' << '\n';
        h << box_str(Box2{{0, 0}, {n - 1, n - 1}, {0, 0}}) << ' ' << '\n' << step << ' ' << '\n';
        h << dx[0] << ' ' << dx[1] << ' ' << '\n' << 0 << '\n' << 0 << '\n';
        h << 0 << ' ' << boxes.size() << ' ' << time << '\n' << step << '\n';
        for (const Box2 &b : boxes)
            for (int d = 0; d < 2; ++d)
                h << lo[d] + b.lo[d] * dx[d] << ' ' << lo[d] + (b.hi[d] + 1) * dx[d] << '\n';
        h << "Level_0/Cell\n";
        if (!h) return amr_fail(ctx, AMRESSIF_ERR_ARG, "write failed: " + root + "/Header");
    }
    std::vector<Field> f;
    for (const double *d : data) f.push_back(Field{d, nullptr, n, n, -1});
    std::string err;
    if (!write_vismf(root + "/Level_0/Cell", boxes, f, err)) return amr_fail(ctx, AMRESSIF_ERR_ARG, err);
    return AMRESSIF_OK;
}

}  // namespace

extern "C" {

int amressif_write_checkpoint(amressif_ctx *ctx, const char *dir, int step, double time)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_write_checkpoint: not available for 3D or periodic inputs");
    if (!ctx || !dir || step < 0) return amr_fail(ctx, AMRESSIF_ERR_ARG, "write_checkpoint: bad argument");
    const amr::Ctx &c = ctx->c;
    const int n = c.g.n;
    std::vector<double> u((size_t)n * (n + 1)), v((size_t)n * (n + 1)), up(u.size()), vp(u.size()),
        p((size_t)n * n);
    if (int rc = amr_state_to_host(ctx, u.data(), v.data(), up.data(), vp.data(), p.data(), nullptr)) return rc;
    const std::string root = std::string(dir) + "/" + step_name("checkpoint", step);
    if (!make_dir(root) || !make_dir(root + "/Level_0"))
        return amr_fail(ctx, AMRESSIF_ERR_ARG, "cannot create " + root + "/Level_0");
    const std::vector<Box2> boxes = max_size_boxes(n, c.in.max_grid_size[0], c.in.max_grid_size[1]);
    {
        std::ofstream h(root + "/Header", std::ios::trunc | std::ios::binary);
        if (!h) return amr_fail(ctx, AMRESSIF_ERR_ARG, "cannot open " + root + "/Header");
        h.precision(15);
        h << "Checkpoint file for AMRESSIF\n" << time << '\n';
        write_boxarray(h, boxes);
        h << '\n';
        if (!h) return amr_fail(ctx, AMRESSIF_ERR_ARG, "write failed: " + root + "/Header");
    }
    struct Item {
        const char *name;
        const double *a;
        int w, nodal;
    } items[] = {{"pressure", p.data(), n, -1},      {"vel_xCont", u.data(), n + 1, 0},
                 {"vel_yCont", v.data(), n, 1},      {"vel_xContPrev", up.data(), n + 1, 0},
                 {"vel_yContPrev", vp.data(), n, 1}};
    std::string err;
    for (const Item &it : items) {
        std::vector<Field> f{Field{it.a, nullptr, it.w, it.nodal == 1 ? n + 1 : n, it.nodal}};
        if (!write_vismf(root + "/Level_0/" + it.name, boxes, f, err)) return amr_fail(ctx, AMRESSIF_ERR_ARG, err);
    }
    return AMRESSIF_OK;
}

int amressif_read_checkpoint(amressif_ctx *ctx, const char *path, double *time)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_read_checkpoint: not available for 3D or periodic inputs");
    if (!ctx || !path) return amr_fail(ctx, AMRESSIF_ERR_ARG, "read_checkpoint: bad argument");
    const int n = ctx->c.g.n;
    const std::string root(path);
    std::string text;
    if (!read_file(root + "/Header", text)) return amr_fail(ctx, AMRESSIF_ERR_ARG, "cannot read " + root + "/Header");
    Scan sc(text);
    sc.line();  // title
    const double t = sc.real();
    const std::vector<Box2> boxes = sc.boxarray();
    if (!sc.ok || boxes.empty()) return amr_fail(ctx, AMRESSIF_ERR_ARG, "malformed " + root + "/Header");
    std::vector<unsigned char> cover((size_t)n * n, 0);
    for (const Box2 &b : boxes) {
        if (b.lo[0] < 0 || b.lo[1] < 0 || b.hi[0] >= n || b.hi[1] >= n || b.hi[0] < b.lo[0] || b.hi[1] < b.lo[1])
            return amr_fail(ctx, AMRESSIF_ERR_ARG, "checkpoint box " + box_str(b) + " is outside the domain");
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i) cover[(size_t)i + (size_t)n * j]++;
    }
    for (unsigned char x : cover)
        if (x != 1) return amr_fail(ctx, AMRESSIF_ERR_ARG, "checkpoint boxes do not tile the domain");
    std::vector<double> u((size_t)n * (n + 1)), v(u.size()), up(u.size()), vp(u.size()), p((size_t)n * n);
    struct Item {
        const char *name;
        double *a;
        int w, nodal;
    } items[] = {{"pressure", p.data(), n, -1},      {"vel_xCont", u.data(), n + 1, 0},
                 {"vel_yCont", v.data(), n, 1},      {"vel_xContPrev", up.data(), n + 1, 0},
                 {"vel_yContPrev", vp.data(), n, 1}};
    std::string err;
    for (const Item &it : items) {
        std::vector<Field> f{Field{nullptr, it.a, it.w, it.nodal == 1 ? n + 1 : n, it.nodal}};
        if (!read_vismf(root + "/Level_0/" + it.name, f, n, err)) return amr_fail(ctx, AMRESSIF_ERR_ARG, err);
    }
    // main.cpp:386-700: the state is velCont, velContPrev and p; phi is not
    // saved (the step's Poisson solve does not read it)
    if (int rc = amr_state_from_host(ctx, u.data(), v.data(), up.data(), vp.data(), p.data(), nullptr)) return rc;
    if (time) *time = t;
    return AMRESSIF_OK;
}

int amressif_write_plotfile(amressif_ctx *ctx, const char *dir, const char *name, int step, double time)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_write_plotfile: not available for 3D or periodic inputs");
    if (!ctx || !dir || !name || step < 0) return amr_fail(ctx, AMRESSIF_ERR_ARG, "write_plotfile: bad argument");
    const amr::Ctx &c = ctx->c;
    const int n = c.g.n;
    std::vector<double> u((size_t)n * (n + 1)), v(u.size()), p((size_t)n * n), U(p.size()), V(p.size());
    if (int rc = amr_state_to_host(ctx, u.data(), v.data(), nullptr, nullptr, p.data(), nullptr)) return rc;
    // velCart of the valid cells: face averages (utilities.cpp:47-54)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            U[(size_t)i + (size_t)n * j] = 0.5 * (u[(size_t)i + (size_t)(n + 1) * j] + u[(size_t)i + 1 + (size_t)(n + 1) * j]);
            V[(size_t)i + (size_t)n * j] = 0.5 * (v[(size_t)i + (size_t)n * j] + v[(size_t)i + (size_t)n * (j + 1)]);
        }
    return write_plot(ctx, dir, name, step, time, {"pressure", "U", "V"}, {p.data(), U.data(), V.data()});
}

int amressif_write_flux_plotfile(amressif_ctx *ctx, const char *dir, int step, double time)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_write_flux_plotfile: not available for 3D or periodic inputs");
    if (!ctx || !dir || step < 0) return amr_fail(ctx, AMRESSIF_ERR_ARG, "write_flux_plotfile: bad argument");
    amr::Ctx &c = ctx->c;
    if (!c.pred_valid)
        return amr_fail(ctx, AMRESSIF_ERR_ARG, "write_flux_plotfile: no step has run since the state was set");
    const int n = c.g.n, P = c.g.P;
    const size_t cells = (size_t)n * n;
    // fluxConvect and fluxViscous of the last step's predictor u* (held in S1)
    // into the first four of the module scratch arrays
    if (int rc = amr_module_scratch(ctx)) return rc;
    double *out[4] = {c.Mc[0], c.Mc[1], c.M[3][0], c.M[3][1]};
    hipError_t e = amr::launch_flux_fields(c, c.S1[0], c.S1[1], nullptr, nullptr, nullptr, nullptr, out[0], out[1],
                                           out[2], out[3]);
    std::vector<double> host(6 * cells, 0.0);  // press_grad_x, press_grad_y stay 0 (fluxPrsGrad, main.cpp:846)
    for (int q = 0; q < 4 && e == hipSuccess; ++q)
        e = hipMemcpy2DAsync(host.data() + q * cells, sizeof(double) * n, out[q], sizeof(double) * P,
                             sizeof(double) * n, n, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) return amr_fail(ctx, AMRESSIF_ERR_HIP, std::string("write_flux_plotfile: ") + hipGetErrorString(e));
    std::vector<const double *> data;
    for (int q = 0; q < 6; ++q) data.push_back(host.data() + q * cells);
    return write_plot(ctx, dir, "pltFlux", step, time,
                      {"conv_flux_x", "conv_flux_y", "visc_flux_x", "visc_flux_y", "press_grad_x", "press_grad_y"},
                      data);
}

int amressif_init_state(amressif_ctx *ctx, const double *psi_coeffs, int kmax)
{
    if (ctx && ctx->c.gen) {
        if (psi_coeffs)
            return amr_fail(ctx, AMRESSIF_ERR_UNSUPPORTED, "init_state: the seeded field is 2D walls only");
        return amr_gen_init_state(ctx);
    }
    if (!ctx || (psi_coeffs && (kmax < 1 || kmax > 64)))
        return amr_fail(ctx, AMRESSIF_ERR_ARG, "init_state: bad argument");
    const int n = ctx->c.g.n;
    std::vector<double> u((size_t)n * (n + 1), 0.0), v(u.size(), 0.0);
    if (psi_coeffs) {
        // psi(x, y) = sum_{kx, ky} a[kx][ky] sin(kx pi x) sin(ky pi y) sin^2(pi x) sin^2(pi y)
        // at the nodes x = i / n, y = j / n (0 on the boundary nodes)
        const double pi = 3.14159265358979323846;
        std::vector<double> S((size_t)kmax * (n + 1)), env(n + 1), T((size_t)kmax * (n + 1)),
            psi((size_t)(n + 1) * (n + 1), 0.0);
        for (int i = 0; i <= n; ++i) {
            const double x = (double)i / n;
            for (int k = 0; k < kmax; ++k) S[(size_t)k * (n + 1) + i] = std::sin(pi * ((double)(k + 1) * x));
            const double s = std::sin(pi * x);
            env[i] = s * s;
        }
        for (int kx = 0; kx < kmax; ++kx)
            for (int j = 0; j <= n; ++j) {
                double t = 0.0;
                for (int ky = 0; ky < kmax; ++ky) t += psi_coeffs[(size_t)kx * kmax + ky] * S[(size_t)ky * (n + 1) + j];
                T[(size_t)kx * (n + 1) + j] = t;
            }
        for (int j = 1; j < n; ++j)
            for (int i = 1; i < n; ++i) {
                double s = 0.0;
                for (int kx = 0; kx < kmax; ++kx) s += S[(size_t)kx * (n + 1) + i] * T[(size_t)kx * (n + 1) + j];
                psi[(size_t)i + (size_t)(n + 1) * j] = s * (env[j] * env[i]);
            }
        const double h = 1.0 / n;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i <= n; ++i)
                u[(size_t)i + (size_t)(n + 1) * j] =
                    (psi[(size_t)i + (size_t)(n + 1) * (j + 1)] - psi[(size_t)i + (size_t)(n + 1) * j]) / h;
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i < n; ++i)
                v[(size_t)i + (size_t)n * j] =
                    -(psi[(size_t)i + 1 + (size_t)(n + 1) * j] - psi[(size_t)i + (size_t)(n + 1) * j]) / h;
    }
    // velContPrev = velCont; p = phi = 0 (kn_init.H:35-36)
    return amr_state_from_host(ctx, u.data(), v.data(), u.data(), v.data(), nullptr, nullptr);
}

}  // extern "C"
