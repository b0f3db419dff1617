// io.cpp — .raw input, VTK image output without the VTK library, checkpoints.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "ls_host.h"

namespace lshost {

std::vector<uint8_t> read_raw(const std::string& path, const Grid& g, int label_bytes,
                              size_t header) {
    if (label_bytes != 1 && label_bytes != 2) throw std::runtime_error("label bytes must be 1 or 2");
    std::ifstream fin(path, std::ios::binary);
    if (!fin.is_open()) throw std::runtime_error("Cannot open file: " + path);
    const size_t bytes = g.n() * (size_t)label_bytes;
    std::vector<uint8_t> out(bytes);
    fin.seekg((std::streamoff)header);
    fin.read((char*)out.data(), (std::streamsize)bytes);
    if ((size_t)fin.gcount() != bytes)
        throw std::runtime_error("short .raw file: " + path + " (" + std::to_string(fin.gcount()) +
                                 " of " + std::to_string(bytes) + " bytes after the header)");
    return out;
}

void write_raw(const std::string& path, const void* data, size_t bytes) {
    std::ofstream fout(path, std::ios::binary);
    if (!fout.is_open()) throw std::runtime_error("Cannot open file for writing: " + path);
    fout.write((const char*)data, (std::streamsize)bytes);
}

namespace {

std::string step_tag(size_t tstep) {
    std::ostringstream s;
    s << std::setw(6) << std::setfill('0') << tstep;
    return s.str();
}

std::string basename_of(const std::string& p) {
    const size_t q = p.find_last_of("/\\");
    return q == std::string::npos ? p : p.substr(q + 1);
}

struct VtkArray {
    const char* name;
    const char* type;
    int comps;
    unsigned flag;
};

// The reference's arrays in its declaration order (src/Simulation.cpp:1323-1345).
const VtkArray kArrays[] = {
    {"MaterialType", nullptr, 1, VTK_TYPE},
    {"Concentration", "Float64", 1, VTK_CONC},
    {"Pressure", "Float64", 1, VTK_PRESSURE},
    {"SulphideFrac", "Float32", 1, VTK_FRAC},
    {"Flux", "Float64", 3, VTK_FLUX_VEC},
    {"Saturation", "Float64", 1, VTK_SATURATION},
    {"CapillaryPressure", "Float64", 1, VTK_CAP_PRESSURE},
    {"RelativePermeability", "Float64", 1, VTK_REL_PERMEABILITY},
    {"EffectivePermeability", "Float64", 1, VTK_PERMEABILITY},
    {"AcidConcentration", "Float64", 1, VTK_ACID_CONC},
};

// Gather one array in VTK point order (i fastest): out[(k*nj + j)*ni + i].
template <typename T, typename F>
void gather(const Grid& g, std::vector<char>& buf, int comps, F get) {
    const size_t n = g.n();
    buf.resize(n * comps * sizeof(T));
    T* o = (T*)buf.data();
    size_t c = 0;
    for (int k = 0; k < g.nk; ++k)
        for (int j = 0; j < g.nj; ++j)
            for (int i = 0; i < g.ni; ++i) {
                const size_t v = ((size_t)i * g.nj + j) * g.nk + k;
                for (int q = 0; q < comps; ++q) o[c++] = (T)get(v, q);
            }
}

}  // namespace

size_t write_vtk(const std::string& root, size_t tstep, HostState& s, unsigned flags) {
    const Grid& g = s.g;
    const std::string tag = step_tag(tstep);
    const char* label_type = s.label_bytes == 2 ? "UInt16" : "UInt8";
    // the reference's scrub of non-finite pressure and capillary pressure
    // before output (src/Simulation.cpp:1487-1517)
    size_t nan_p = 0;
    for (double& x : s.pressure)
        if (!std::isfinite(x)) { x = 0.0; ++nan_p; }
    for (double& x : s.pc)
        if (!std::isfinite(x)) x = 0.0;

    // master .pvti
    {
        std::ofstream f(root + "_" + tag + ".pvti");
        if (!f.is_open()) throw std::runtime_error("cannot write " + root + "_" + tag + ".pvti");
        f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PImageData\" version=\"0.1\">\n";
        f << "\t<PImageData WholeExtent=\"0 " << g.ni - 1 << " 0 " << g.nj - 1 << " 0 " << g.nk - 1
          << "\" GhostLevel=\"0\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n\t\t<PPointData>\n";
        for (const VtkArray& a : kArrays) {
            if (!(flags & a.flag)) continue;
            f << "\t\t\t<PDataArray type=\"" << (a.type ? a.type : label_type) << "\" Name=\""
              << a.name << "\"";
            if (a.comps > 1) f << " NumberOfComponents=\"" << a.comps << "\"";
            f << "/>\n";
        }
        f << "\t\t</PPointData>\n";
        f << "\t\t<Piece Extent=\"0 " << g.ni - 1 << " 0 " << g.nj - 1 << " 0 " << g.nk - 1
          << "\" Source=\"" << basename_of(root) << "_0_" << tag << ".vti\"/>\n";
        f << "\t</PImageData>\n</VTKFile>\n";
    }
    // the piece: appended raw binary, one UInt64 byte count before each block
    std::vector<std::vector<char>> blocks;
    std::ostringstream head;
    head << "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"1.0\" "
            "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    head << "  <ImageData WholeExtent=\"0 " << g.ni - 1 << " 0 " << g.nj - 1 << " 0 " << g.nk - 1
         << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n";
    head << "    <Piece Extent=\"0 " << g.ni - 1 << " 0 " << g.nj - 1 << " 0 " << g.nk - 1 << "\">\n";
    head << "      <PointData" << ((flags & VTK_FLUX_VEC) ? " Vectors=\"Flux\"" : "") << ">\n";
    uint64_t offset = 0;
    for (const VtkArray& a : kArrays) {
        if (!(flags & a.flag)) continue;
        std::vector<char> b;
        switch (a.flag) {
            case VTK_TYPE:
                if (s.label_bytes == 2)
                    gather<uint16_t>(g, b, 1, [&](size_t v, int) { return ((const uint16_t*)s.img.data())[v]; });
                else
                    gather<uint8_t>(g, b, 1, [&](size_t v, int) { return s.img[v]; });
                break;
            case VTK_CONC: gather<double>(g, b, 1, [&](size_t v, int) { return s.conc[v]; }); break;
            case VTK_PRESSURE: gather<double>(g, b, 1, [&](size_t v, int) { return s.pressure[v]; }); break;
            case VTK_FRAC: gather<float>(g, b, 1, [&](size_t v, int) { return s.frac[v]; }); break;
            case VTK_FLUX_VEC:
                gather<double>(g, b, 3, [&](size_t v, int q) {
                    return q == 0 ? s.qx[v] : q == 1 ? s.qy[v] : s.qz[v];
                });
                break;
            case VTK_SATURATION: gather<double>(g, b, 1, [&](size_t v, int) { return s.saturation[v]; }); break;
            case VTK_CAP_PRESSURE: gather<double>(g, b, 1, [&](size_t v, int) { return s.pc[v]; }); break;
            case VTK_REL_PERMEABILITY: gather<double>(g, b, 1, [&](size_t v, int) { return s.krel[v]; }); break;
            case VTK_PERMEABILITY: gather<double>(g, b, 1, [&](size_t v, int) { return s.perm[v]; }); break;
            default: gather<double>(g, b, 1, [&](size_t v, int) { return s.acid[v]; }); break;
        }
        head << "        <DataArray type=\"" << (a.type ? a.type : label_type) << "\" Name=\"" << a.name
             << "\"";
        if (a.comps > 1) head << " NumberOfComponents=\"" << a.comps << "\"";
        head << " format=\"appended\" offset=\"" << offset << "\"/>\n";
        offset += sizeof(uint64_t) + b.size();
        blocks.push_back(std::move(b));
    }
    head << "      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n  </ImageData>\n";
    head << "  <AppendedData encoding=\"raw\">\n   _";
    const std::string vti = root + "_0_" + tag + ".vti";
    std::ofstream f(vti, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("cannot write " + vti);
    const std::string h = head.str();
    f.write(h.data(), (std::streamsize)h.size());
    for (const auto& b : blocks) {
        const uint64_t nb = b.size();
        f.write((const char*)&nb, sizeof(nb));
        f.write(b.data(), (std::streamsize)b.size());
    }
    const std::string tail = "\n  </AppendedData>\n</VTKFile>\n";
    f.write(tail.data(), (std::streamsize)tail.size());
    return nan_p;
}

void write_pvd(const std::string& path, const std::vector<double>& t,
               const std::vector<size_t>& steps, const std::string& root_basename) {
    // src/main.cpp:510-526
    std::ofstream pvd(path);
    if (!pvd.is_open()) throw std::runtime_error("cannot write " + path);
    pvd << "<?xml version=\"1.0\"?>\n";
    pvd << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    pvd << "  <Collection>\n";
    for (size_t q = 0; q < t.size(); ++q)
        pvd << "    <DataSet timestep=\"" << t[q] << "\" file=\"" << root_basename << "_"
            << step_tag(steps[q]) << ".pvti\"/>\n";
    pvd << "  </Collection>\n</VTKFile>\n";
}

namespace {
const char kMagic[8] = {'L', 'S', 'C', 'H', 'K', 'P', 'T', '1'};
constexpr int32_t kVersion = 1;

struct CkHeader {
    char magic[8];
    int32_t version, ni, nj, nk, label_bytes, pad;
    int64_t step;
    double t, dt;
};

template <typename T>
void put(std::ofstream& f, const std::vector<T>& v) {
    f.write((const char*)v.data(), (std::streamsize)(v.size() * sizeof(T)));
}
template <typename T>
void get(std::ifstream& f, std::vector<T>& v, size_t n) {
    v.resize(n);
    f.read((char*)v.data(), (std::streamsize)(n * sizeof(T)));
    if ((size_t)f.gcount() != n * sizeof(T)) throw std::runtime_error("checkpoint file is short");
}
}  // namespace

void write_checkpoint(const std::string& dir, int step, double t, double dt, const HostState& s) {
    const std::string path = dir + "/checkpoint_" + step_tag((size_t)step) + ".lsck";
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("cannot write checkpoint " + path);
    CkHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.ni = s.g.ni; h.nj = s.g.nj; h.nk = s.g.nk;
    h.label_bytes = s.label_bytes;
    h.step = step; h.t = t; h.dt = dt;
    f.write((const char*)&h, sizeof(h));
    put(f, s.img); put(f, s.frac); put(f, s.inventory); put(f, s.saturation);
    put(f, s.pressure); put(f, s.conc); put(f, s.acid);
    if (!f) throw std::runtime_error("checkpoint write failed: " + path);
}

bool load_checkpoint(const std::string& dir, int step, HostState& s, double* t, double* dt) {
    const size_t n = s.g.n();
    const std::string own = dir + "/checkpoint_" + step_tag((size_t)step) + ".lsck";
    std::ifstream f(own, std::ios::binary);
    if (f.is_open()) {
        CkHeader h{};
        f.read((char*)&h, sizeof(h));
        if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion)
            throw std::runtime_error("not a checkpoint of this format: " + own);
        if (h.ni != s.g.ni || h.nj != s.g.nj || h.nk != s.g.nk || h.label_bytes != s.label_bytes)
            throw std::runtime_error("checkpoint grid or label size differs from the run: " + own);
        get(f, s.img, n * (size_t)s.label_bytes); get(f, s.frac, n); get(f, s.inventory, n);
        get(f, s.saturation, n); get(f, s.pressure, n); get(f, s.conc, n); get(f, s.acid, n);
        if (t) *t = h.t;
        if (dt) *dt = h.dt;
        return true;
    }
    // the reference's one-rank file (no ghost planes on a single rank)
    const std::string ref = dir + "/checkpoint_rank0_" + step_tag((size_t)step) + ".bin";
    std::ifstream r(ref, std::ios::binary | std::ios::ate);
    if (!r.is_open())
        throw std::runtime_error("failed to locate checkpoint file: " + own + " or " + ref);
    const size_t expect = n * ((size_t)s.label_bytes + sizeof(float) + 5 * sizeof(double));
    if ((size_t)r.tellg() != expect)
        throw std::runtime_error("reference checkpoint size differs from a one-rank run of this grid: " + ref);
    r.seekg(0);
    get(r, s.img, n * (size_t)s.label_bytes); get(r, s.frac, n); get(r, s.inventory, n);
    get(r, s.saturation, n); get(r, s.pressure, n); get(r, s.conc, n); get(r, s.acid, n);
    return false;
}

}  // namespace lshost
