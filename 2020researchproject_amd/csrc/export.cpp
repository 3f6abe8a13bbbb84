// export.cpp - netCDF classic (CDF-1) result, level-set and restart files.
//
// The reference links libnetcdf (src/ibm3d_export.c:16-650,
// src/ibm3d_init.c:45-224); this library writes and reads the same classic
// format itself: dimensions time = 1, x = Nx+4, y = Ny+4, z = Nz+4; double
// coordinates of the padded cell centres; float fields [time][z][y][x] with
// NaN in solid cells. The restart is lossy exactly like the reference's:
// fields round-trip through float32 and the face fluxes are re-interpolated
// from the cell values on the next IBMSolver_iterate.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "solver.h"

namespace ibm {

namespace {

enum { NC_CHAR_T = 2, NC_INT_T = 4, NC_FLOAT_T = 5, NC_DOUBLE_T = 6 };

struct NcVar {
    std::string name;
    std::vector<int> dims;
    std::vector<std::pair<std::string, std::string>> atts;  // text attributes
    int type;
    std::vector<char> data;  // big-endian payload
};

void put_u32(std::vector<char> &b, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((char)((v >> s) & 0xff));
}
void put_name(std::vector<char> &b, const std::string &s) {
    put_u32(b, (uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    while (b.size() % 4) b.push_back(0);
}
template <typename T>
void append_be(std::vector<char> &out, const T *v, size_t n) {
    const size_t w = sizeof(T);
    const size_t o = out.size();
    out.resize(o + n * w);
    for (size_t i = 0; i < n; i++) {
        unsigned char tmp[8];
        memcpy(tmp, &v[i], w);
        for (size_t k = 0; k < w; k++) out[o + i * w + k] = (char)tmp[w - 1 - k];
    }
}

int type_size(int t) { return t == NC_DOUBLE_T ? 8 : t == NC_CHAR_T ? 1 : 4; }

// Classic format with 32-bit offsets (CDF-1), or with 64-bit offsets (CDF-2)
// when the file would pass the 2 GiB that CDF-1 offsets address. Returns
// false (nothing written) when the file cannot be opened or a variable is
// too large for the format.
bool write_cdf1(const std::string &path, const std::vector<std::pair<std::string, int>> &dims,
                const std::vector<NcVar> &vars) {
    size_t payload = 0;
    for (auto &v : vars) {
        if (v.data.size() > 0xFFFFFFFCull) {
            fprintf(stderr, "error: variable %s (%zu bytes) exceeds the netCDF classic variable size\n",
                    v.name.c_str(), v.data.size());
            return false;
        }
        payload += (v.data.size() + 3) / 4 * 4;
    }
    // header size does not depend on the offset values, only on their width
    auto header = [&](int version, std::vector<size_t> *begin_pos) {
        std::vector<char> h;
        h.insert(h.end(), {'C', 'D', 'F', (char)version});
        put_u32(h, 0);  // numrecs: no record dimension
        put_u32(h, 0x0A);
        put_u32(h, (uint32_t)dims.size());
        for (auto &d : dims) { put_name(h, d.first); put_u32(h, (uint32_t)d.second); }
        put_u32(h, 0); put_u32(h, 0);  // no global attributes
        put_u32(h, 0x0B);
        put_u32(h, (uint32_t)vars.size());
        for (auto &v : vars) {
            put_name(h, v.name);
            put_u32(h, (uint32_t)v.dims.size());
            for (int d : v.dims) put_u32(h, (uint32_t)d);
            if (v.atts.empty()) { put_u32(h, 0); put_u32(h, 0); }
            else {
                put_u32(h, 0x0C);
                put_u32(h, (uint32_t)v.atts.size());
                for (auto &a : v.atts) {
                    put_name(h, a.first);
                    put_u32(h, NC_CHAR_T);
                    put_u32(h, (uint32_t)a.second.size());
                    h.insert(h.end(), a.second.begin(), a.second.end());
                    while (h.size() % 4) h.push_back(0);
                }
            }
            put_u32(h, (uint32_t)v.type);
            put_u32(h, (uint32_t)((v.data.size() + 3) / 4 * 4));
            begin_pos->push_back(h.size());
            put_u32(h, 0);  // begin, patched below
            if (version == 2) put_u32(h, 0);
        }
        return h;
    };
    std::vector<size_t> begin_pos;
    std::vector<char> h = header(1, &begin_pos);
    const int version = h.size() + payload > 0x7FFFFFFFull ? 2 : 1;
    if (version == 2) { begin_pos.clear(); h = header(2, &begin_pos); }
    size_t off = h.size();
    for (size_t q = 0; q < vars.size(); q++) {
        const int w = version == 2 ? 8 : 4;
        for (int s = 0; s < w; s++) h[begin_pos[q] + s] = (char)((off >> (8 * (w - 1 - s))) & 0xff);
        off += (vars[q].data.size() + 3) / 4 * 4;
    }
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    fwrite(h.data(), 1, h.size(), f);
    for (auto &v : vars) {
        fwrite(v.data.data(), 1, v.data.size(), f);
        static const char pad[4] = {0, 0, 0, 0};
        fwrite(pad, 1, (4 - v.data.size() % 4) % 4, f);
    }
    fclose(f);
    return true;
}

NcVar var_d(const std::string &name, std::vector<int> dims, const double *v, size_t n,
            std::vector<std::pair<std::string, std::string>> atts) {
    NcVar x{name, std::move(dims), std::move(atts), NC_DOUBLE_T, {}};
    append_be(x.data, v, n);
    return x;
}
NcVar var_f(const std::string &name, std::vector<int> dims, const float *v, size_t n,
            std::vector<std::pair<std::string, std::string>> atts) {
    NcVar x{name, std::move(dims), std::move(atts), NC_FLOAT_T, {}};
    append_be(x.data, v, n);
    return x;
}
NcVar var_i(const std::string &name, std::vector<int> dims, const int *v, size_t n) {
    NcVar x{name, std::move(dims), {}, NC_INT_T, {}};
    append_be(x.data, v, n);
    return x;
}

// minimal CDF-1/CDF-2 reader: variable name -> (type, dimension lengths,
// raw big-endian bytes). Every read is bounded by the buffer; a malformed or
// truncated header makes parse() return false.
struct NcFile {
    std::vector<char> buf;
    struct V { int type; size_t begin, vsize; std::vector<int> dims; };
    std::vector<std::pair<std::string, V>> vars;
    std::vector<int> dimlen;
    size_t pos = 0;
    int version = 1;
    bool bad = false;
    bool have(size_t n) {
        if (pos > buf.size() || buf.size() - pos < n) bad = true;
        return !bad;
    }
    uint32_t u32() {
        if (!have(4)) return 0;
        uint32_t v = 0;
        for (int s = 0; s < 4; s++) v = (v << 8) | (unsigned char)buf[pos++];
        return v;
    }
    uint64_t u64() { uint64_t a = u32(); return (a << 32) | u32(); }
    std::string name() {
        const uint32_t n = u32();
        const size_t padded = ((size_t)n + 3) / 4 * 4;
        if (!have(padded)) return std::string();
        std::string s(buf.data() + pos, n);
        pos += padded;
        return s;
    }
    bool parse() {
        if (buf.size() < 32 || buf[0] != 'C' || buf[1] != 'D' || buf[2] != 'F') return false;
        version = buf[3];
        if (version != 1 && version != 2) return false;
        pos = 4;
        u32();  // numrecs
        uint32_t tag = u32(), n = u32();
        if (tag == 0x0A)
            for (uint32_t q = 0; q < n && !bad; q++) { name(); dimlen.push_back((int)u32()); }
        tag = u32(); n = u32();
        if (tag == 0x0C) skip_atts(n);
        tag = u32(); n = u32();
        if (bad) return false;
        if (tag != 0x0B) return n == 0;
        for (uint32_t q = 0; q < n && !bad; q++) {
            V v;
            const std::string nm = name();
            const uint32_t nd = u32();
            for (uint32_t d = 0; d < nd && !bad; d++) {
                const uint32_t id = u32();
                if (id >= dimlen.size()) bad = true;
                else v.dims.push_back(dimlen[id]);
            }
            const uint32_t at = u32(), an = u32();
            if (at == 0x0C) skip_atts(an);
            v.type = (int)u32();
            v.vsize = u32();
            v.begin = version == 2 ? (size_t)u64() : (size_t)u32();
            if (v.begin > buf.size() || buf.size() - v.begin < v.vsize) bad = true;
            vars.push_back({nm, v});
        }
        return !bad;
    }
    void skip_atts(uint32_t n) {
        for (uint32_t a = 0; a < n && !bad; a++) {
            name();
            const int t = (int)u32();
            const uint32_t cnt = u32();
            const size_t len = ((size_t)cnt * type_size(t) + 3) / 4 * 4;
            if (have(len)) pos += len;
        }
    }
    const V *find(const std::string &nm) const {
        for (auto &v : vars) if (v.first == nm) return &v.second;
        return nullptr;
    }
    // n values of v (the caller checked v's type and shape; parse() checked
    // that its vsize bytes lie inside the buffer)
    template <typename T>
    void read(const V &v, T *out, size_t n) const {
        const size_t w = sizeof(T);
        if (n * w > v.vsize) fatal("netCDF variable shorter than its shape");
        for (size_t i = 0; i < n; i++) {
            unsigned char tmp[8];
            for (size_t k = 0; k < w; k++) tmp[k] = (unsigned char)buf[v.begin + i * w + w - 1 - k];
            memcpy(&out[i], tmp, w);
        }
    }
};

std::vector<double> fetch_ref(IBMSolver *s, int field) {
    std::vector<double> r((size_t)(s->Nx + 4) * (s->Ny + 4) * (s->Nz + 4));
    IBMSolver_get_field(s, field, r.data());
    return r;
}

// reference layout (k fastest) -> float [z][y][x]
std::vector<float> to_zyx(const IBMSolver *s, const std::vector<double> &ref) {
    const int X = s->Nx + 4, Y = s->Ny + 4, Z = s->Nz + 4;
    std::vector<float> out((size_t)X * Y * Z);
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++)
            for (int k = 0; k < Z; k++) out[((size_t)k * Y + j) * X + i] = (float)ref[((size_t)i * Y + j) * Z + k];
    return out;
}

std::string with_ext(const char *filename) { return std::string(filename) + ".nc"; }

}  // namespace

// <- src/ibm3d_export.c:16-253
void export_result_impl(IBMSolver *s, const char *filename) {
    if (!s->assembled) fatal("IBMSolver_assemble must precede IBMSolver_export_*");
    const int X = s->Nx + 4, Y = s->Ny + 4, Z = s->Nz + 4;
    const std::vector<float> u = to_zyx(s, fetch_ref(s, IBM3D_FIELD_U1));
    const std::vector<float> v = to_zyx(s, fetch_ref(s, IBM3D_FIELD_U2));
    const std::vector<float> w = to_zyx(s, fetch_ref(s, IBM3D_FIELD_U3));
    const std::vector<float> p = to_zyx(s, fetch_ref(s, IBM3D_FIELD_P));
    const size_t n = (size_t)X * Y * Z;
    std::vector<float> vort(n), Q(n);
    const double *xv = s->xc.data(), *yv = s->yc.data(), *zv = s->zc.data();
    auto at = [&](const std::vector<float> &a, int k, int j, int i) { return a[((size_t)k * Y + j) * X + i]; };
    // vorticity magnitude and Q by central differences clamped at the array
    // ends (:111-141)
    for (int k = 0; k < Z; k++)
        for (int j = 0; j < Y; j++)
            for (int i = 0; i < X; i++) {
                const int ip = std::max(i - 1, 0), in = std::min(i + 1, X - 1);
                const int jp = std::max(j - 1, 0), jn = std::min(j + 1, Y - 1);
                const int kp = std::max(k - 1, 0), kn = std::min(k + 1, Z - 1);
                const float dudx = (at(u, k, j, in) - at(u, k, j, ip)) / (xv[in] - xv[ip]);
                const float dudy = (at(u, k, jn, i) - at(u, k, jp, i)) / (yv[jn] - yv[jp]);
                const float dudz = (at(u, kn, j, i) - at(u, kp, j, i)) / (zv[kn] - zv[kp]);
                const float dvdx = (at(v, k, j, in) - at(v, k, j, ip)) / (xv[in] - xv[ip]);
                const float dvdy = (at(v, k, jn, i) - at(v, k, jp, i)) / (yv[jn] - yv[jp]);
                const float dvdz = (at(v, kn, j, i) - at(v, kp, j, i)) / (zv[kn] - zv[kp]);
                const float dwdx = (at(w, k, j, in) - at(w, k, j, ip)) / (xv[in] - xv[ip]);
                const float dwdy = (at(w, k, jn, i) - at(w, k, jp, i)) / (yv[jn] - yv[jp]);
                const float dwdz = (at(w, kn, j, i) - at(w, kp, j, i)) / (zv[kn] - zv[kp]);
                const float ox = dwdy - dvdz, oy = dudz - dwdx, oz = dvdx - dudy;
                const size_t q = ((size_t)k * Y + j) * X + i;
                vort[q] = std::sqrt(ox * ox + oy * oy + oz * oz);
                Q[q] = dudx * dvdy + dvdy * dwdz + dwdz * dudx - dudy * dvdx - dvdz * dwdy - dudz * dwdx;
            }
    const std::vector<std::pair<std::string, int>> dims = {{"time", 1}, {"x", X}, {"y", Y}, {"z", Z}};
    const double t = s->time;
    const int it = s->iter;
    std::vector<NcVar> vars;
    vars.push_back(var_d("time", {0}, &t, 1, {{"axis", "T"}, {"units", "seconds"}}));
    vars.push_back(var_d("x", {1}, xv, X, {{"axis", "X"}, {"units", "meters"}}));
    vars.push_back(var_d("y", {2}, yv, Y, {{"axis", "Y"}, {"units", "meters"}}));
    vars.push_back(var_d("z", {3}, zv, Z, {{"axis", "Z"}, {"units", "meters"}}));
    vars.push_back(var_i("iter", {}, &it, 1));
    vars.push_back(var_f("u", {0, 3, 2, 1}, u.data(), n, {{"units", "m/s"}}));
    vars.push_back(var_f("v", {0, 3, 2, 1}, v.data(), n, {{"units", "m/s"}}));
    vars.push_back(var_f("w", {0, 3, 2, 1}, w.data(), n, {{"units", "m/s"}}));
    vars.push_back(var_f("p", {0, 3, 2, 1}, p.data(), n, {{"units", "m/s"}}));
    vars.push_back(var_f("vorticity", {0, 3, 2, 1}, vort.data(), n, {{"units", "1/s"}}));
    vars.push_back(var_f("Q", {0, 3, 2, 1}, Q.data(), n, {{"units", "1/s^2"}}));
    if (!write_cdf1(with_ext(filename), dims, vars))
        fprintf(stderr, "error: cannot open file %s\n", with_ext(filename).c_str());  // :145-147 (continue)
}

// <- src/ibm3d_export.c:262-435
void export_lvset_flag_impl(IBMSolver *s, const char *filename) {
    if (!s->assembled) fatal("IBMSolver_assemble must precede IBMSolver_export_*");
    const int X = s->Nx + 4, Y = s->Ny + 4, Z = s->Nz + 4;
    const size_t n = (size_t)X * Y * Z;
    const std::vector<float> lv = to_zyx(s, s->lvset_h);
    std::vector<int> fl(n);
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++)
            for (int k = 0; k < Z; k++) fl[((size_t)k * Y + j) * X + i] = s->flag_ref[((size_t)i * Y + j) * Z + k];
    const std::vector<std::pair<std::string, int>> dims = {{"x", X}, {"y", Y}, {"z", Z}};
    std::vector<NcVar> vars;
    vars.push_back(var_d("x", {0}, s->xc.data(), X, {{"axis", "X"}, {"units", "meters"}}));
    vars.push_back(var_d("y", {1}, s->yc.data(), Y, {{"axis", "Y"}, {"units", "meters"}}));
    vars.push_back(var_d("z", {2}, s->zc.data(), Z, {{"axis", "Z"}, {"units", "meters"}}));
    vars.push_back(var_f("lvset", {2, 1, 0}, lv.data(), n, {}));
    vars.push_back(var_i("flag", {2, 1, 0}, fl.data(), n));
    if (!write_cdf1(with_ext(filename), dims, vars))
        fprintf(stderr, "error: cannot open file %s\n", with_ext(filename).c_str());
}

// <- src/ibm3d_export.c:437-650 (u~ formed from u* and p on cells -1..N)
void export_intermediate_impl(IBMSolver *s, const char *filename) {
    if (!s->assembled) fatal("IBMSolver_assemble must precede IBMSolver_export_*");
    const int X = s->Nx + 4, Y = s->Ny + 4, Z = s->Nz + 4;
    const size_t n = (size_t)X * Y * Z;
    std::vector<double> us[3] = {fetch_ref(s, IBM3D_FIELD_U1_STAR), fetch_ref(s, IBM3D_FIELD_U2_STAR),
                                 fetch_ref(s, IBM3D_FIELD_U3_STAR)};
    const std::vector<double> p = fetch_ref(s, IBM3D_FIELD_P), pp = fetch_ref(s, IBM3D_FIELD_P_PRIME);
    std::vector<double> ut[3];
    const std::vector<double> *c[3] = {&s->xc, &s->yc, &s->zc};
    const long st[3] = {(long)Y * Z, Z, 1};
    for (int a = 0; a < 3; a++) {
        ut[a].assign(n, 0.0);
        for (int i = -1; i <= s->Nx; i++)
            for (int j = -1; j <= s->Ny; j++)
                for (int k = -1; k <= s->Nz; k++) {
                    const long q = host_ref_index(s, i, j, k);
                    const int id = a == 0 ? i : a == 1 ? j : k;
                    ut[a][q] = us[a][q] + s->dt * (p[q + st[a]] - p[q - st[a]]) / ((*c[a])[id + 3] - (*c[a])[id + 1]);
                }
    }
    const std::vector<std::pair<std::string, int>> dims = {{"time", 1}, {"x", X}, {"y", Y}, {"z", Z}};
    const double t = s->time;
    const int it = s->iter;
    std::vector<NcVar> vars;
    vars.push_back(var_d("time", {0}, &t, 1, {{"axis", "T"}, {"units", "seconds"}}));
    vars.push_back(var_d("x", {1}, s->xc.data(), X, {{"axis", "X"}, {"units", "meters"}}));
    vars.push_back(var_d("y", {2}, s->yc.data(), Y, {{"axis", "Y"}, {"units", "meters"}}));
    vars.push_back(var_d("z", {3}, s->zc.data(), Z, {{"axis", "Z"}, {"units", "meters"}}));
    vars.push_back(var_i("iter", {}, &it, 1));
    const char *names[7] = {"u_star", "v_star", "w_star", "u_tilde", "v_tilde", "w_tilde", "p_prime"};
    const std::vector<double> *src[7] = {&us[0], &us[1], &us[2], &ut[0], &ut[1], &ut[2], &pp};
    for (int q = 0; q < 7; q++) {
        const std::vector<float> f = to_zyx(s, *src[q]);
        vars.push_back(var_f(names[q], {0, 3, 2, 1}, f.data(), n, {{"units", "m/s"}}));
    }
    if (!write_cdf1(with_ext(filename), dims, vars))
        fprintf(stderr, "error: cannot open file %s\n", with_ext(filename).c_str());
}

// <- src/ibm3d_init.c:45-224: time, iter, u, v, w, p from "<filename>.nc".
// The file is opened, parsed and checked (types, and the padded grid shape of
// the solver's IBMSolver_set_grid) before anything else, so a missing,
// truncated or mismatched file aborts with its own message.
void init_flow_file_impl(IBMSolver *s, const char *filename) {
    NcFile nc;
    FILE *f = fopen(with_ext(filename).c_str(), "rb");
    if (!f) {
        case_error(std::string("error: cannot open file ") + filename);
    }
    fseek(f, 0, SEEK_END);
    const long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    nc.buf.resize(len > 0 ? len : 0);
    if (len <= 0 || fread(nc.buf.data(), 1, len, f) != (size_t)len) { fclose(f); fatal("cannot read restart file"); }
    fclose(f);
    if (!nc.parse()) fatal("restart file is not netCDF classic, or is truncated");
    const char *need[6] = {"time", "iter", "u", "v", "w", "p"};
    for (const char *nm : need)
        if (!nc.find(nm)) {
            case_error(std::string("error: varible '") + nm + "' not found");
        }
    const int X = s->Nx + 4, Y = s->Ny + 4, Z = s->Nz + 4;
    const NcFile::V &vt = *nc.find("time"), &vi = *nc.find("iter");
    if (vt.type != NC_DOUBLE_T || vi.type != NC_INT_T) fatal("restart file: time must be double and iter int");
    const std::vector<int> shape = {1, Z, Y, X};
    const char *fields[4] = {"u", "v", "w", "p"};
    for (const char *nm : fields) {
        const NcFile::V &v = *nc.find(nm);
        if (v.type != NC_FLOAT_T || v.dims != shape) {
            char buf[200];
            snprintf(buf, sizeof buf, "error: restart variable '%s' must be float [1][%d][%d][%d] (the solver's "
                     "padded grid)", nm, Z, Y, X);
            case_error(buf);
        }
    }
    // the file is checked in full before the solver is touched
    if (!s->assembled) fatal("IBMSolver_assemble must precede IBMSolver_init_flow_file");
    s->conv_ready = false;
    nc.read(vt, &s->time, 1);
    nc.read(vi, &s->iter, 1);
    const size_t n = (size_t)X * Y * Z;
    const int ids[4] = {IBM3D_FIELD_U1, IBM3D_FIELD_U2, IBM3D_FIELD_U3, IBM3D_FIELD_P};
    std::vector<float> buf(n);
    std::vector<double> ref(n);
    for (int a = 0; a < 4; a++) {
        nc.read(*nc.find(fields[a]), buf.data(), n);
        for (int i = 0; i < X; i++)
            for (int j = 0; j < Y; j++)
                for (int k = 0; k < Z; k++) ref[((size_t)i * Y + j) * Z + k] = buf[((size_t)k * Y + j) * X + i];
        IBMSolver_set_field(s, ids[a], ref.data());
    }
}

}  // namespace ibm
