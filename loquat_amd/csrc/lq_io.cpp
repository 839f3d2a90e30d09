// lq_io.cpp -- parameters and project files.
//
//   lq_params_derive    InitializeParameters   src/loquat.cu:111-139 (+ dt, :331-333)
//   lq_load_parameters  LoquatIo::LoadParameters src/loquat_io.cu:69-201
//   lq_load_particles   LoquatIo::LoadParticles  src/loquat_io.cu:204-237
//   lq_save_output      LoquatIo::SaveXml + SaveParticles  src/loquat_io.cu:240-549
//   lq_load_restart     (FP64 positions and void ratio of an output; no counterpart)
//
// The XML is the reference's flat list of <key type="int|float">value</key>
// elements under one root; a small reader replaces tinyxml2 + Boost.Variant.
// Errors come back as status codes with a message (lq_last_error) instead of
// exit(EXIT_FAILURE).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/loquat_amd.h"

namespace lq {
void set_error(const std::string& msg);
}

namespace {

const float kPi = 3.14159265358979323846f;  // PI as a float (include/types.h:30)

struct XmlValue {
    bool is_int;
    int i;
    float f;
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) ++a;
    while (b > a && isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Parses the children of the root element into key -> typed value.
lq_status parse_flat_xml(const std::string& text, std::map<std::string, XmlValue>& out) {
    size_t pos = 0;
    const size_t n = text.size();
    int depth = 0;
    while (pos < n) {
        size_t lt = text.find('<', pos);
        if (lt == std::string::npos) break;
        if (text.compare(lt, 4, "<!--") == 0) {
            size_t e = text.find("-->", lt + 4);
            if (e == std::string::npos) {
                lq::set_error("unterminated XML comment");
                return LQ_ERR_PARSE;
            }
            pos = e + 3;
            continue;
        }
        if (text.compare(lt, 2, "<?") == 0) {
            size_t e = text.find("?>", lt + 2);
            if (e == std::string::npos) {
                lq::set_error("unterminated XML declaration");
                return LQ_ERR_PARSE;
            }
            pos = e + 2;
            continue;
        }
        size_t gt = text.find('>', lt);
        if (gt == std::string::npos) {
            lq::set_error("unterminated XML tag");
            return LQ_ERR_PARSE;
        }
        std::string tag = text.substr(lt + 1, gt - lt - 1);
        if (!tag.empty() && tag[0] == '/') {
            --depth;
            pos = gt + 1;
            continue;
        }
        bool self_closing = !tag.empty() && tag.back() == '/';
        if (self_closing) tag.pop_back();
        std::string name;
        size_t i = 0;
        while (i < tag.size() && !isspace((unsigned char)tag[i])) name += tag[i++];
        std::string attrs = tag.substr(i);
        if (depth == 0) {  // root element
            if (!self_closing) ++depth;
            pos = gt + 1;
            continue;
        }
        // child element: <name type="...">value</name>
        std::string close = "</" + name + ">";
        size_t ce = self_closing ? gt : text.find(close, gt + 1);
        if (ce == std::string::npos) {
            lq::set_error("element <" + name + "> is not closed");
            return LQ_ERR_PARSE;
        }
        std::string value = self_closing ? std::string() : trim(text.substr(gt + 1, ce - gt - 1));
        std::string type;
        size_t tp = attrs.find("type");
        if (tp != std::string::npos) {
            size_t q1 = attrs.find_first_of("\"'", tp);
            if (q1 != std::string::npos) {
                size_t q2 = attrs.find(attrs[q1], q1 + 1);
                if (q2 != std::string::npos) type = attrs.substr(q1 + 1, q2 - q1 - 1);
            }
        }
        XmlValue v{};
        if (type == "int") {
            v.is_int = true;
            v.i = atoi(value.c_str());
        } else if (type == "float") {
            v.is_int = false;
            v.f = (float)atof(value.c_str());
        } else {
            lq::set_error("Wrong data type for " + name);  // src/loquat_io.cu:106-110
            return LQ_ERR_PARSE;
        }
        out[name] = v;
        pos = self_closing ? gt + 1 : ce + close.size();
    }
    return LQ_OK;
}

struct Keys {
    const std::map<std::string, XmlValue>& m;
    lq_status st = LQ_OK;
    // CheckParameter (src/loquat_io.cu:56-66): a missing key is an error.
    const XmlValue* get(const char* key, bool want_int) {
        auto it = m.find(key);
        if (it == m.end()) {
            if (st == LQ_OK) {
                lq::set_error(std::string(key) + " does not exist in the xml file");
                st = LQ_ERR_MISSING_KEY;
            }
            return nullptr;
        }
        if (it->second.is_int != want_int) {
            if (st == LQ_OK) {
                lq::set_error(std::string("wrong value type for ") + key);
                st = LQ_ERR_PARSE;
            }
            return nullptr;
        }
        return &it->second;
    }
    int i(const char* key) {
        const XmlValue* v = get(key, true);
        return v ? v->i : 0;
    }
    float f(const char* key) {
        const XmlValue* v = get(key, false);
        return v ? v->f : 0.0f;
    }
};

}  // namespace

extern "C" lq_status lq_params_derive(lq_params* p) {
    if (!p) {
        lq::set_error("lq_params_derive: null params");
        return LQ_ERR_INVALID_ARG;
    }
    if (!(p->dr > 0.0f) || !(p->hdr > 0.0f) || !(p->cs > 0.0f) || !(p->rho0 > 0.0f)) {
        lq::set_error("particle-size, h-over-dr, speed-of-sound and reference-density must be positive");
        return LQ_ERR_INVALID_ARG;
    }
    p->h = p->dr * p->hdr;
    p->i_h = 1.0f / p->h;
    p->eta = 0.01f * p->h * p->h;
    p->m = p->dr * p->dr * p->dr * p->rho0;
    p->cell_num_x = static_cast<int>(0.5f * p->i_h * p->domain_size_x) + 1;
    p->cell_num_y = static_cast<int>(0.5f * p->i_h * p->domain_size_y) + 1;
    p->cell_num_z = static_cast<int>(0.5f * p->i_h * p->domain_size_z) + 1;
    const int64_t nol = (int64_t)p->cell_num_x * p->cell_num_y * p->cell_num_z;
    if (p->cell_num_x <= 0 || p->cell_num_y <= 0 || p->cell_num_z <= 0 || 2 * nol + 1 > (int64_t(1) << 31) - 1) {
        lq::set_error("cell grid is empty or exceeds 2^31 cells");
        return LQ_ERR_INVALID_ARG;
    }
    p->ncell_one_layer = (int)nol;
    p->ncell = 2 * p->ncell_one_layer + 1;
    p->ncell_delete = p->ncell - 1;
    // Wendland C2 constants (src/loquat.cu:129-131).
    p->kernel_normalization_par = 42.0f / 256.0f / (powf(p->h, 3) * kPi);
    p->kernel_gradient_normalization_par = -5.0f * 42.0f / 256.0f / (powf(p->h, 5) * kPi);
    if (p->artificial_pressure == 1) {
        p->kernel_zero = p->kernel_normalization_par * (2.0f) * (2.0f) * (2.0f) * (2.0f) * (0.5f);
        const float q = p->dr * p->i_h;
        p->kernel_dr = p->kernel_normalization_par * (2.0f - q) * (2.0f - q) * (2.0f - q) * (2.0f - q) * (q + 0.5f);
    }
    // Constant time step (src/loquat.cu:325-333: the dynamic branch is disabled).
    p->dt = p->cfl * p->h / p->cs;
    return LQ_OK;
}

extern "C" lq_status lq_load_parameters(const char* prefix, lq_params* par, lq_vw_params* vw) {
    if (!prefix || !par) {
        lq::set_error("lq_load_parameters: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    std::string name = std::string(prefix) + ".xml";
    std::ifstream in(name.c_str(), std::ios::binary);
    if (!in) {
        lq::set_error("Cannot load the parameter file: " + name);
        return LQ_ERR_IO;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    std::map<std::string, XmlValue> kv;
    lq_status st = parse_flat_xml(ss.str(), kv);
    if (st != LQ_OK) return st;
    if (kv.empty()) {
        lq::set_error("no parameters in " + name);
        return LQ_ERR_PARSE;
    }
    Keys k{kv};
    lq_params p;
    std::memset(&p, 0, sizeof(p));
    p.np = k.i("number-of-particles");
    p.npb = k.i("number-of-boundary-particles");
    p.npm = k.i("number-of-material-particles");
    if (k.st == LQ_OK && p.np != p.npb + p.npm) {
        lq::set_error("Wrong number of particles: np != npb + npm");  // src/loquat_io.cu:127-131
        return LQ_ERR_INVALID_ARG;
    }
    p.domain_size_x = k.f("domain-size-x");
    p.domain_size_y = k.f("domain-size-y");
    p.domain_size_z = k.f("domain-size-z");
    p.domain_min_x = k.f("domain-min-x");
    p.domain_min_y = k.f("domain-min-y");
    p.domain_min_z = k.f("domain-min-z");
    p.dr = k.f("particle-size");
    p.hdr = k.f("h-over-dr");
    p.rho0 = k.f("reference-density");
    p.cs = k.f("speed-of-sound");
    p.cfl = k.f("cfl-number");
    p.constitutive_model = k.i("constitutive-model");
    if (p.constitutive_model == 1) {
        p.model_par_1 = k.f("elastic-K");
        p.model_par_2 = k.f("elastic-G");
        p.model_par_3 = k.f("DP-kphi");
        p.model_par_4 = k.f("DP-kpsi");
        p.model_par_5 = k.f("DP-kcohesion");
    } else if (p.constitutive_model == 2) {
        p.model_par_1 = k.f("hypo-par-c1");
        p.model_par_2 = k.f("hypo-par-c2");
        p.model_par_3 = k.f("hypo-par-c3");
        p.model_par_4 = k.f("hypo-par-c4");
        p.model_par_5 = k.f("cohesion");
    } else if (p.constitutive_model == 3) {
        lq_vw_params v;
        v.phi_c = k.f("vw-phi-c");
        v.hs = k.f("vw-hs");
        v.n = k.f("vw-n");
        v.ed0 = k.f("vw-ed0");
        v.ec0 = k.f("vw-ec0");
        v.ei0 = k.f("vw-ei0");
        v.alpha = k.f("vw-alpha");
        v.beta = k.f("vw-beta");
        if (vw) *vw = v;
    }
    p.acc_x = k.f("acceleration-x");
    p.acc_y = k.f("acceleration-y");
    p.acc_z = k.f("acceleration-z");
    p.time_max = k.f("time-max");
    p.time = k.f("time");
    p.dt = k.f("time-step");
    p.output_frequency = k.f("output-frequency");
    p.output_number = k.i("output-number");
    p.density_shepard_filter = k.i("density-shepard-filter");
    p.stress_shepard_filter = k.i("stress-shepard-filter");
    p.dynamic_time_step = k.i("dynamic-time-step");
    p.integration_method = k.i("integration-method");
    p.viscosity_type = k.i("viscosity-type");
    if (p.viscosity_type == 1) p.artificial_viscosity_alpha = k.f("artificial-viscosity-alpha");
    p.artificial_pressure = k.i("artificial-pressure");
    if (p.artificial_pressure == 1) p.artificial_pressure_coefficient = k.f("artificial-pressure-coefficient");
    p.output_format = k.i("output-format");
    if (k.st != LQ_OK) return k.st;
    *par = p;
    return LQ_OK;
}

extern "C" lq_status lq_load_particles(const char* prefix, const lq_params* par, lq_host_particles* out) {
    if (!prefix || !par || !out || !out->pospres || !out->velrhop || !out->idep || !out->str || !out->stn) {
        lq::set_error("lq_load_particles: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->np < 0) {
        lq::set_error("lq_load_particles: negative particle count");
        return LQ_ERR_INVALID_ARG;
    }
    std::string name = std::string(prefix) + ".dat";
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) {
        lq::set_error("Cannot open " + name);  // src/loquat_io.cu:230-234
        return LQ_ERR_IO;
    }
    std::fseek(f, 0, SEEK_END);
    long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<char> buf((size_t)(size > 0 ? size : 0) + 1, '\0');
    size_t got = size > 0 ? std::fread(buf.data(), 1, (size_t)size, f) : 0;
    std::fclose(f);
    buf[got] = '\0';
    char* c = buf.data();
    // Column order of src/loquat_io.cu:226: pospres.xyzw, velrhop.xyzw,
    // idep.x, idep.y, str (6), stn (6).
    for (int i = 0; i < par->np; ++i) {
        float v[22];
        for (int t = 0; t < 22; ++t) {
            char* e = nullptr;
            errno = 0;
            v[t] = std::strtof(c, &e);
            if (e == c) {
                lq::set_error(name + ": expected 22 numbers per particle, row " + std::to_string(i) +
                              " column " + std::to_string(t) + " unreadable");
                return LQ_ERR_PARSE;
            }
            c = e;
        }
        float* pp = out->pospres + 4 * (size_t)i;
        float* vr = out->velrhop + 4 * (size_t)i;
        float* id = out->idep + 4 * (size_t)i;
        float* s = out->str + 6 * (size_t)i;
        float* n = out->stn + 6 * (size_t)i;
        for (int t = 0; t < 4; ++t) pp[t] = v[t];
        for (int t = 0; t < 4; ++t) vr[t] = v[4 + t];
        id[0] = v[8];
        id[1] = v[9];
        id[2] = 0.0f;
        id[3] = 0.0f;
        for (int t = 0; t < 6; ++t) s[t] = v[10 + t];
        for (int t = 0; t < 6; ++t) n[t] = v[16 + t];
    }
    return LQ_OK;
}

// ------------------------------------------------------------------ output
// LoquatIo::SaveXml / SaveParticles (src/loquat_io.cu:240-549), written so
// that every output is a restart: see lq_save_output in the header for the
// differences from the reference.
namespace {

// Shortest form that reads back as the same float (9 significant digits).
struct F9 {
    float v;
};
std::ostream& operator<<(std::ostream& o, F9 x) {
    char b[32];
    std::snprintf(b, sizeof(b), "%.9g", (double)x.v);
    return o << b;
}

bool write_xml(const std::string& name, const lq_params& p, const lq_vw_params* vw) {
    std::ofstream o(name.c_str(), std::ofstream::out);
    if (!o) return false;
    auto fl = [&](const char* k, float v) { o << "  <" << k << " type=\"float\">" << F9{v} << "</" << k << ">\n"; };
    auto in = [&](const char* k, int v) { o << "  <" << k << " type=\"int\">" << v << "</" << k << ">\n"; };
    o << "<?xml version=\"1.0\" encoding=\"utf-8\" ?>\n<parameters>\n";
    in("number-of-particles", p.np);
    in("number-of-boundary-particles", p.npb);
    in("number-of-material-particles", p.npm);
    fl("domain-size-x", p.domain_size_x);
    fl("domain-size-y", p.domain_size_y);
    fl("domain-size-z", p.domain_size_z);
    fl("domain-min-x", p.domain_min_x);
    fl("domain-min-y", p.domain_min_y);
    fl("domain-min-z", p.domain_min_z);
    fl("particle-size", p.dr);
    fl("h-over-dr", p.hdr);
    fl("reference-density", p.rho0);
    fl("speed-of-sound", p.cs);
    fl("cfl-number", p.cfl);
    in("constitutive-model", p.constitutive_model);
    if (p.constitutive_model == 1) {
        fl("elastic-K", p.model_par_1);
        fl("elastic-G", p.model_par_2);  // the reader's key (the reference writes poisson-G)
        fl("DP-kphi", p.model_par_3);
        fl("DP-kpsi", p.model_par_4);
        fl("DP-kcohesion", p.model_par_5);
    } else if (p.constitutive_model == 2) {
        fl("hypo-par-c1", p.model_par_1);
        fl("hypo-par-c2", p.model_par_2);
        fl("hypo-par-c3", p.model_par_3);
        fl("hypo-par-c4", p.model_par_4);
        fl("cohesion", p.model_par_5);
    } else if (p.constitutive_model == 3 && vw) {
        fl("vw-phi-c", vw->phi_c);
        fl("vw-hs", vw->hs);
        fl("vw-n", vw->n);
        fl("vw-ed0", vw->ed0);
        fl("vw-ec0", vw->ec0);
        fl("vw-ei0", vw->ei0);
        fl("vw-alpha", vw->alpha);
        fl("vw-beta", vw->beta);
    }
    fl("acceleration-x", p.acc_x);
    fl("acceleration-y", p.acc_y);
    fl("acceleration-z", p.acc_z);
    fl("time-max", p.time_max);
    fl("time", p.time);
    fl("time-step", p.dt);
    fl("output-frequency", p.output_frequency);
    in("output-number", p.output_number);
    in("density-shepard-filter", p.density_shepard_filter);
    in("stress-shepard-filter", p.stress_shepard_filter);
    in("dynamic-time-step", p.dynamic_time_step);
    in("integration-method", p.integration_method);
    in("viscosity-type", p.viscosity_type);
    if (p.viscosity_type == 1) fl("artificial-viscosity-alpha", p.artificial_viscosity_alpha);
    in("artificial-pressure", p.artificial_pressure);
    if (p.artificial_pressure == 1) fl("artificial-pressure-coefficient", p.artificial_pressure_coefficient);
    in("output-format", p.output_format);
    o << "</parameters>\n";
    return (bool)o;
}

bool write_dat(const std::string& name, int np, const lq_host_particles& h) {
    std::ofstream o(name.c_str(), std::ofstream::out);
    if (!o) return false;
    for (int i = 0; i < np; ++i) {
        const float* pp = h.pospres + 4 * (size_t)i;
        const float* vr = h.velrhop + 4 * (size_t)i;
        const float* id = h.idep + 4 * (size_t)i;
        const float* s = h.str + 6 * (size_t)i;
        const float* n = h.stn + 6 * (size_t)i;
        o << F9{pp[0]} << ' ' << F9{pp[1]} << ' ' << F9{pp[2]} << ' ' << F9{pp[3]} << ' ' << F9{vr[0]} << ' '
          << F9{vr[1]} << ' ' << F9{vr[2]} << ' ' << F9{vr[3]} << ' ' << F9{id[0]} << ' ' << F9{id[1]};
        for (int t = 0; t < 6; ++t) o << ' ' << F9{s[t]};
        for (int t = 0; t < 6; ++t) o << ' ' << F9{n[t]};
        o << '\n';
    }
    return (bool)o;
}

// Binary sidecar with what the 22-column .dat cannot hold: FP64 positions
// and the void ratio (idep.z) of constitutive model 3.
const char kRstMagic[8] = {'L', 'Q', 'R', 'S', 'T', '0', '0', '1'};

bool write_rst(const std::string& name, int np, const lq_host_particles& h, const double* posd) {
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) return false;
    const int64_t n = np;
    bool ok = std::fwrite(kRstMagic, 1, 8, f) == 8 && std::fwrite(&n, 8, 1, f) == 1;
    std::vector<double> xyz(3 * (size_t)np);
    for (int i = 0; i < np; ++i)
        for (int a = 0; a < 3; ++a)
            xyz[3 * (size_t)i + a] = posd ? posd[3 * (size_t)i + a] : (double)h.pospres[4 * (size_t)i + a];
    std::vector<float> e((size_t)np);
    for (int i = 0; i < np; ++i) e[i] = h.idep[4 * (size_t)i + 2];
    ok = ok && (np == 0 || std::fwrite(xyz.data(), 8, xyz.size(), f) == xyz.size());
    ok = ok && (np == 0 || std::fwrite(e.data(), 4, e.size(), f) == e.size());
    ok = (std::fclose(f) == 0) && ok;
    return ok;
}

// One .vtu per object id (idep.x), objects in descending id order and
// particles in input order inside each (src/loquat_io.cu:373-545).
bool write_vtu(const std::string& dir, int number, int np, const lq_host_particles& h) {
    std::vector<int> idx((size_t)np);
    for (int i = 0; i < np; ++i) idx[i] = i;
    auto id_of = [&](int i) { return (int)h.idep[4 * (size_t)i]; };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return id_of(a) > id_of(b); });
    size_t s0 = 0;
    while (s0 < idx.size()) {
        const int id = id_of(idx[s0]);
        size_t s1 = s0;
        while (s1 < idx.size() && id_of(idx[s1]) == id) ++s1;
        const size_t cnt = s1 - s0;
        std::ostringstream nm;
        nm << dir << "/" << id << "_" << number << ".vtu";
        std::ofstream o(nm.str().c_str(), std::ofstream::out);
        if (!o) return false;
        auto arr = [&](const char* name, int comps, auto&& put) {
            o << "        <DataArray type=\"Float32\" Name=\"" << name << "\"";
            if (comps > 1) o << " NumberOfComponents=\"" << comps << "\"";
            o << " format=\"ascii\">\n";
            for (size_t r = s0; r < s1; ++r) put(idx[r]);
            o << "</DataArray>\n";
        };
        o << "<?xml version=\"1.0\"?>\n"
             "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
             "  <UnstructuredGrid GhostLevel=\"0\">\n"
          << "    <Piece NumberOfPoints=\"" << cnt << "\" NumberOfCells=\"" << cnt << "\" Name=\"" << id << "\">\n"
          << "      <PointData>\n";
        arr("Pressure", 1, [&](int i) { o << F9{h.pospres[4 * (size_t)i + 3]} << ' '; });
        arr("Velocity", 3, [&](int i) {
            for (int a = 0; a < 3; ++a) o << F9{h.velrhop[4 * (size_t)i + a]} << ' ';
        });
        arr("Density", 1, [&](int i) { o << F9{h.velrhop[4 * (size_t)i + 3]} << ' '; });
        arr("PlasticStrain", 1, [&](int i) { o << F9{h.idep[4 * (size_t)i + 1]} << ' '; });
        arr("Stress", 6, [&](int i) {
            for (int a = 0; a < 6; ++a) o << F9{h.str[6 * (size_t)i + a]} << ' ';
        });
        arr("Strain", 6, [&](int i) {
            for (int a = 0; a < 6; ++a) o << F9{h.stn[6 * (size_t)i + a]} << ' ';
        });
        o << "      </PointData>\n      <CellData>\n      </CellData>\n      <Points>\n";
        arr("Points", 3, [&](int i) {
            for (int a = 0; a < 3; ++a) o << F9{h.pospres[4 * (size_t)i + a]} << ' ';
        });
        o << "      </Points>\n      <Cells>\n"
             "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
        for (size_t j = 0; j < cnt; ++j) o << j << ' ';
        o << "</DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
        for (size_t j = 0; j < cnt; ++j) o << j + 1 << ' ';
        o << "</DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
        for (size_t j = 0; j < cnt; ++j) o << "1 ";
        o << "</DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
        if (!o) return false;
        s0 = s1;
    }
    return true;
}

}  // namespace

extern "C" lq_status lq_save_output(const char* dir, const lq_params* par, const lq_vw_params* vw,
                                    const lq_host_particles* h, const double* posd) {
    if (!dir || !par || !h || !h->pospres || !h->velrhop || !h->idep || !h->str || !h->stn) {
        lq::set_error("lq_save_output: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->np < 0) {
        lq::set_error("lq_save_output: negative particle count");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->constitutive_model == 3 && !vw) {
        lq::set_error("lq_save_output: constitutive model 3 needs lq_vw_params");
        return LQ_ERR_INVALID_ARG;
    }
    const std::string d(dir);
    const std::string n = std::to_string(par->output_number);
    if (!write_xml(d + "/output_" + n + ".xml", *par, vw)) {
        lq::set_error("Cannot write " + d + "/output_" + n + ".xml");
        return LQ_ERR_IO;
    }
    if (!write_dat(d + "/" + n + ".dat", par->np, *h) || !write_rst(d + "/" + n + ".rst", par->np, *h, posd)) {
        lq::set_error("Cannot write " + d + "/" + n + ".dat / .rst");
        return LQ_ERR_IO;
    }
    if (par->output_format == 2 && !write_vtu(d, par->output_number, par->np, *h)) {
        lq::set_error("Cannot write the .vtu files in " + d);
        return LQ_ERR_IO;
    }
    return LQ_OK;
}

extern "C" lq_status lq_load_restart(const char* prefix, const lq_params* par, double* posd, float* idep) {
    if (!prefix || !par || !posd) {
        lq::set_error("lq_load_restart: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    const std::string name = std::string(prefix) + ".rst";
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) {
        lq::set_error("Cannot open " + name);
        return LQ_ERR_IO;
    }
    char magic[8];
    int64_t n = -1;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kRstMagic, 8) == 0 &&
              std::fread(&n, 8, 1, f) == 1 && n == par->np;
    std::vector<float> e((size_t)std::max<int64_t>(n, 0));
    ok = ok && (n == 0 || std::fread(posd, 8, 3 * (size_t)n, f) == 3 * (size_t)n);
    ok = ok && (n == 0 || std::fread(e.data(), 4, e.size(), f) == e.size());
    std::fclose(f);
    if (!ok) {
        lq::set_error(name + ": not a restart file for " + std::to_string(par->np) + " particles");
        return LQ_ERR_PARSE;
    }
    if (idep)
        for (int64_t i = 0; i < n; ++i) idep[4 * (size_t)i + 2] = e[i];
    return LQ_OK;
}
