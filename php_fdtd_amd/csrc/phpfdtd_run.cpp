// phpfdtd_run.cpp - reference-compatible driver of libphpfdtd_hip.
//
// Reads a PhP_FDTD case the way fdtd.c's main() does on one rank - the cfg
// keys (read_config / parse_source / parse_material, fdtd.c:1403-1920), the
// built-in material library (fdtd.c:626-715), OUTCAR / QeZV / QeZM / DMM /
// dielectric-table files with switchxz (fdtd.c:346-552, 2248-2470) and
// medium_map.bin (fdtd.c:718-959) - steps the whole domain on the GPU through
// the C ABI (include/phpfdtd.h) in place of the CPU loops, and writes the
// reference's outputs: source.txt, field_slice_<axis><pos>_<step>.bin, the
// monitor file, the full-monitor file and the progress / divergence lines
// (fdtd.c:3102-3917).  Outputs at step t hold the fields after that step's
// update, as the reference's do.
//
//   phpfdtd_run case.cfg [--fp32 | --fp32-rv64] [--precise]
//
// --fp32 selects the precision form (ion r, v stored as compensated FP32),
// --fp32-rv64 the precision form with r, v stored in FP64; --precise writes
// the monitor files with %.17e instead of the reference's %e.
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "../../include/phpfdtd.h"

namespace {

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) ++a;
    while (b > a && isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// strtok-style split: empty fields between consecutive delimiters are dropped
std::vector<std::string> split(const std::string &s, char d) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : s) {
        if (c == d) {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

// " key = value" of one comma-separated token: the key up to '=', the value
// the first blank-delimited word after it (the reference's sscanf pattern).
bool key_word(const std::string &tok, std::string *key, std::string *val) {
    const size_t eq = tok.find('=');
    if (eq == std::string::npos) return false;
    *key = trim(tok.substr(0, eq));
    size_t p = eq + 1;
    while (p < tok.size() && isspace((unsigned char)tok[p])) ++p;
    size_t e = p;
    while (e < tok.size() && !isspace((unsigned char)tok[e])) ++e;
    if (e == p || key->empty()) return false;
    *val = trim(tok.substr(p, e - p));
    return true;
}

struct Mat {
    std::string name = "Vacuum";
    double re[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, im[9] = {0};
    double sigma = 0;
    int natom = 0, use_ade = 0, switchxz = 0;
    double anglez = 0, anglex = 0;
    std::string born_file, qezv_file, qezm_file, dmm_file, diel_file;
    bool born_eps = false, qezv_set = false, qezm_set = false, dmm_set = false, use_diel = false;
    std::vector<double> qezv, qezm, dmm;
};

struct Cfg {
    int nx = 50, ny = 50, nz = 50, tsteps = 500;
    double dx = 0.01, dy = 0.01, dz = 0.01;
    int pml = 10;
    double pml_r = 0.001;
    int out_int = 20;
    std::string medium_file = "medium_map.bin";
    double c = 3e8, eps0 = 8.854e-12, mu0 = M_PI * 4e-7;
    std::vector<phpfdtd_source> src;
    double dt = -1.0;
    std::string out_comp = "Ex";
    char axis = 'z';
    int pos = 25;
    bool mon = false;
    int mx = 0, my = 0, mz = 0, matom = 0, mdir = 0;
    std::string mon_file;
    bool full = false;
    std::string full_file = "full_monitor.dat";
    double fsrc = -1.0;
    std::vector<Mat> mats;
};

[[noreturn]] void die(const char *fmt, const char *a = nullptr) {
    fprintf(stderr, fmt, a);
    fprintf(stderr, "\n");
    exit(1);
}

// parse_source (fdtd.c:1403-1496): defaults are taken from the configuration
// as it stands when the line is read (the first pass, before nx or dt).
void parse_source(Cfg &c, const std::string &line) {
    phpfdtd_source s;
    memset(&s, 0, sizeof(s));
    s.type = 0;
    s.x = c.nx / 2;
    s.y = c.ny / 2;
    s.z = c.nz / 3;
    s.amplitude = 1.0;
    s.frequency = 1e9;
    s.t0 = 40 * c.dt;
    s.tau = 15 * c.dt;
    s.component = 'z';
    s.start_time = 0.0;
    s.end_time = c.tsteps * c.dt;
    bool has_f = false, has_wn = false;
    double wn = 0;
    for (const auto &tok : split(line, ',')) {
        std::string k, v;
        if (!key_word(tok, &k, &v)) continue;
        if (k == "type") {
            if (v == "gaussian") s.type = 0;
            else if (v == "sinusoidal") s.type = 1;
            else if (v == "ricker") s.type = 2;
            else if (v == "dipole") s.type = 3;
        } else if (k == "x") s.x = atoi(v.c_str());
        else if (k == "y") s.y = atoi(v.c_str());
        else if (k == "z") s.z = atoi(v.c_str());
        else if (k == "amplitude") s.amplitude = strtod(v.c_str(), nullptr);
        else if (k == "frequency") { s.frequency = strtod(v.c_str(), nullptr); has_f = true; }
        else if (k == "wavenumber") { wn = strtod(v.c_str(), nullptr); has_wn = true; }
        else if (k == "t0") s.t0 = strtod(v.c_str(), nullptr);
        else if (k == "tau") s.tau = strtod(v.c_str(), nullptr);
        else if (k == "component") s.component = v[0];
        else if (k == "start_time") s.start_time = strtod(v.c_str(), nullptr);
        else if (k == "end_time") s.end_time = strtod(v.c_str(), nullptr);
    }
    if (s.type == 3) {  // only dipole sources convert a wavenumber (fdtd.c:1491-1506)
        if (has_f && has_wn) die("Error: Cannot specify both frequency and wavenumber for dipole source");
        if (!has_f && !has_wn) die("Error: Must specify either frequency or wavenumber for dipole source");
        if (has_wn) s.frequency = wn * 0.03e12;
    }
    c.src.push_back(s);
}

// Built-in materials 0-3 at the source frequency (fdtd.c:626-715).
void builtin_materials(Cfg &c) {
    c.mats.assign(4, Mat());
    const double w = c.fsrc / 0.03e12;  // cm^-1
    Mat &si = c.mats[1];
    si.name = "SiO2";
    const double e0 = 1.5, s1 = 231, g1 = 69, w1 = 806, s2 = 866, g2 = 75, w2 = 1063;
    const double t1 = w1 * w1 - w * w, t2 = w2 * w2 - w * w;
    const double er = e0 + s1 * s1 * t1 / (t1 * t1 + w * w * g1 * g1) + s2 * s2 * t2 / (t2 * t2 + w * w * g2 * g2);
    const double ei = s1 * s1 * w * g1 / (t1 * t1 + w * w * g1 * g1) + s2 * s2 * w * g2 / (t2 * t2 + w * w * g2 * g2);
    for (int d = 0; d < 3; ++d) {
        si.re[4 * d] = er;
        si.im[4 * d] = ei;
    }
    if (er <= 0) si.use_ade = 1;
    Mat &bi = c.mats[2];
    bi.name = "Uniaxial Crystal";
    bi.re[0] = 8.0; bi.re[4] = 7.0; bi.re[8] = 5.0;
    Mat &lo = c.mats[3];
    lo.name = "Lossy Medium";
    lo.re[0] = lo.re[4] = lo.re[8] = 9.0;
    lo.sigma = 0.1;
}

// parse_material (fdtd.c:1521-1727)
void parse_material(Cfg &c, const std::string &line) {
    int id = -1;
    for (const auto &tok : split(line, ',')) {
        std::string k, v;
        if (key_word(tok, &k, &v) && k == "id") {
            id = atoi(v.c_str());
            break;
        }
    }
    if (id < 0) {
        fprintf(stderr, "Error: Material must have a valid ID\n");
        return;
    }
    if (id >= (int)c.mats.size()) c.mats.resize(id + 1);
    Mat &m = c.mats[id];
    static const char *ek[9] = {"xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz"};
    for (const auto &tok : split(line, ',')) {
        std::string k, v;
        if (!key_word(tok, &k, &v)) continue;
        bool done = false;
        for (int e = 0; e < 9 && !done; ++e) {
            if (k == std::string("eps_") + ek[e]) { m.re[e] = atof(v.c_str()); done = true; }
            else if (k == std::string("eps_") + ek[e] + "i") { m.im[e] = atof(v.c_str()); done = true; }
        }
        if (done) continue;
        if (k == "name") m.name = v;
        else if (k == "sigma") m.sigma = atof(v.c_str());
        else if (k == "dielectric") { m.diel_file = v; m.use_diel = true; }
        else if (k == "switchxz") m.switchxz = atoi(v.c_str());
        else if (k == "natom") m.natom = atoi(v.c_str());
        else if (k == "QeZV") { m.qezv_file = v; m.qezv_set = true; }
        else if (k == "QeZM") { m.qezm_file = v; m.qezm_set = true; }
        else if (k == "DMM") { m.dmm_file = v; m.dmm_set = true; }
        else if (k == "use_ADE") {
            m.use_ade = atoi(v.c_str());
            if (m.use_ade != 0 && m.use_ade != 1) {
                fprintf(stderr, "Error: use_ADE must be 0 or 1\n");
                return;
            }
        } else if (k == "anglez") m.anglez = atof(v.c_str());
        else if (k == "anglex") m.anglex = atof(v.c_str());
        else if (k == "born") {
            // a file name; the inline [..] form cannot survive the comma split
            if (v.find('[') == std::string::npos && v.find(']') == std::string::npos) {
                m.born_file = v;
                m.born_eps = true;
            }
        }
    }
}

// read_config (fdtd.c:1730-1920): sources in a first pass, then the material
// library at the first source's frequency, then every other key.
Cfg read_config(const char *path) {
    Cfg c;
    c.pos = c.nz / 2;
    FILE *f = fopen(path, "r");
    if (!f) {
        perror("Error opening config file");
        exit(1);
    }
    std::vector<std::string> lines;
    char buf[4096];
    while (fgets(buf, sizeof(buf), f)) {
        std::string l(buf);
        const size_t nl = l.find('\n');
        if (nl != std::string::npos) l.resize(nl);
        lines.push_back(l);
    }
    fclose(f);
    for (const auto &l : lines) {
        if (l.empty() || l[0] == '#') continue;
        if (l.compare(0, 6, "source") == 0) {
            parse_source(c, l.substr(6));
            c.fsrc = c.src[0].frequency;
        }
    }
    builtin_materials(c);
    for (const auto &l : lines) {
        if (l.empty() || l[0] == '#') continue;
        if (l.compare(0, 8, "material") == 0) {
            parse_material(c, l.substr(8));
            continue;
        }
        const size_t eq = l.find('=');
        if (eq == std::string::npos || eq == 0) continue;
        const std::string k = trim(l.substr(0, eq)), v = trim(l.substr(eq + 1));
        if (v.empty()) continue;
        if (k == "nx") c.nx = atoi(v.c_str());
        else if (k == "ny") c.ny = atoi(v.c_str());
        else if (k == "nz") c.nz = atoi(v.c_str());
        else if (k == "tsteps") c.tsteps = atoi(v.c_str());
        else if (k == "delta_x") c.dx = atof(v.c_str());
        else if (k == "delta_y") c.dy = atof(v.c_str());
        else if (k == "delta_z") c.dz = atof(v.c_str());
        else if (k == "pml_thickness") c.pml = atoi(v.c_str());
        else if (k == "pml_reflection") c.pml_r = atof(v.c_str());
        else if (k == "output_interval") c.out_int = atoi(v.c_str());
        else if (k == "medium_file") c.medium_file = v;
        else if (k == "dt") c.dt = atof(v.c_str());
        else if (k == "output_component") {
            c.out_comp = v.substr(0, 9);
            static const char *ok[] = {"Ex", "Ey", "Ez", "E", "Dx", "Dy", "Dz"};
            bool good = false;
            for (const char *o : ok) good |= c.out_comp == o;
            if (!good) {
                fprintf(stderr, "output_component must be one of Ex, Ey, Ez, Dx, Dy, Dz, E ('%s'); using 'Ex'\n",
                        c.out_comp.c_str());
                c.out_comp = "Ex";
            }
        } else if (k == "output_slice_axis") {
            c.axis = v[0];
            if (c.axis != 'x' && c.axis != 'y' && c.axis != 'z') {
                fprintf(stderr, "output_slice_axis must be 'x', 'y' or 'z' ('%c'); using 'z'\n", c.axis);
                c.axis = 'z';
            }
        } else if (k == "output_slice_position") c.pos = atoi(v.c_str());
        else if (k == "monitor_point") {
            c.mon = true;
            for (const auto &tok : split(v, ',')) {
                std::string sk, sv;
                if (!key_word(tok, &sk, &sv)) continue;
                if (sk == "x") c.mx = atoi(sv.c_str());
                else if (sk == "y") c.my = atoi(sv.c_str());
                else if (sk == "z") c.mz = atoi(sv.c_str());
                else if (sk == "atom") c.matom = atoi(sv.c_str());
                else if (sk == "dir") {
                    if (sv == "x") c.mdir = 0;
                    else if (sv == "y") c.mdir = 1;
                    else if (sv == "z") c.mdir = 2;
                    else {
                        fprintf(stderr, "Warning: Invalid direction '%s' for monitor_point. Using 'x'.\n", sv.c_str());
                        c.mdir = 0;
                    }
                }
            }
        } else if (k == "monitor_output_file") c.mon_file = v;
        else if (k == "full_monitor") c.full = atoi(v.c_str()) != 0;
        else if (k == "full_monitor_output_file") c.full_file = v;
    }
    double md = c.dx;
    if (c.dy < md) md = c.dy;
    if (c.dz < md) md = c.dz;
    const double cfl = md / (c.c * sqrt(3.0)) * 0.99;
    if (c.dt < 0 || c.dt > cfl) c.dt = cfl;
    for (auto &s : c.src)
        if (s.end_time <= 0) s.end_time = c.tsteps * c.dt;
    return c;
}

// x <-> z exchange of one 3x3 block, M' = P M P with P the x/z permutation
// (the swaps of fdtd.c:2270-2286, 2291-2316, 2333-2349, 2369-2384, 2405-2423).
void switch_xz(double *b, int stride) {
    auto at = [&](int i, int j) -> double & { return b[i * stride + j]; };
    std::swap(at(0, 0), at(2, 2));
    std::swap(at(0, 1), at(2, 1));
    std::swap(at(0, 2), at(2, 0));
    std::swap(at(1, 0), at(1, 2));
}

std::vector<double> read_numbers(const std::string &file, int n) {
    FILE *f = fopen(file.c_str(), "r");
    if (!f) die("Error opening %s", file.c_str());
    printf("%s %d\n", file.c_str(), n);
    std::vector<double> v(n);
    for (int i = 0; i < n; ++i)
        if (fscanf(f, "%lf", &v[i]) != 1) {
            fclose(f);
            die("Error: not enough data in %s", file.c_str());
        }
    fclose(f);
    return v;
}

// Static dielectric tensor from an OUTCAR fragment (read_born_from_file,
// fdtd.c:423-552; only the dielectric block is used by the time step).
bool read_outcar_eps(const std::string &file, int natom, double re[9]) {
    FILE *f = fopen(file.c_str(), "r");
    if (!f) {
        fprintf(stderr, "Error opening Born file: %s\n", file.c_str());
        return false;
    }
    char line[512];
    bool in_diel = false, in_born = false;
    int nd = 0, nb = 0, ion = -1, dir = 0;
    double vals[9];
    while (fgets(line, sizeof(line), f)) {
        line[strcspn(line, "\n")] = 0;
        if (!line[0]) continue;
        if (strstr(line, "MACROSCOPIC STATIC DIELECTRIC TENSOR")) {
            in_diel = true;
            nd = 0;
            continue;
        }
        if (strstr(line, "BORN EFFECTIVE CHARGES")) {
            in_born = true;
            nb = 0;
            continue;
        }
        if ((in_diel || in_born) && strstr(line, "---")) {
            if (in_born) ion = -1;
            continue;
        }
        if (in_diel && nd < 9) {
            double a, b, cc;
            if (sscanf(line, "%*s %lf %lf %lf", &a, &b, &cc) == 3 || sscanf(line, "%lf %lf %lf", &a, &b, &cc) == 3) {
                vals[nd] = a;
                vals[nd + 1] = b;
                vals[nd + 2] = cc;
                nd += 3;
            }
            continue;
        }
        if (in_born) {
            if (const char *p = strstr(line, "ion")) {
                p += 3;
                while (*p && isspace((unsigned char)*p)) ++p;
                ion = atoi(p);
                dir = 0;
                continue;
            }
            int idx;
            double a, b, cc;
            if (sscanf(line, "%d %lf %lf %lf", &idx, &a, &b, &cc) == 4 && ion >= 1 && ion <= natom && dir < 3) {
                nb += 3;
                ++dir;
            }
        }
    }
    fclose(f);
    if (nd < 9 || nb < 9 * natom) {
        fprintf(stderr, "Error: Only %d values read (expected %d) from %s\n", nd + nb, 9 + 9 * natom, file.c_str());
        return false;
    }
    for (int e = 0; e < 9; ++e) re[e] = vals[e];
    return true;
}

// read_dielectric_from_file (fdtd.c:346-401): the row nearest f, symmetric fill.
bool read_dielectric(const std::string &file, double f, double re[9], double im[9]) {
    FILE *fp = fopen(file.c_str(), "r");
    if (!fp) {
        fprintf(stderr, "Error opening dielectric file: %s\n", file.c_str());
        return false;
    }
    char line[256];
    double best = 1e308;
    bool found = false;
    while (fgets(line, sizeof(line), fp)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        double v[15];
        if (sscanf(line, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3],
                   &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11], &v[12], &v[13], &v[14]) != 15)
            continue;
        const double d = fabs(v[0] - f);
        if (d < best) {
            best = d;
            const double r6[6] = {v[3], v[4], v[5], v[6], v[7], v[8]}, i6[6] = {v[9], v[10], v[11], v[12], v[13], v[14]};
            const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};  // xx xy xz / yx yy yz / zx zy zz from xx xy xz yy yz zz
            for (int e = 0; e < 9; ++e) {
                re[e] = r6[map[e]];
                im[e] = i6[map[e]];
            }
            found = true;
        }
    }
    fclose(fp);
    if (!found) {
        fprintf(stderr, "Error: No valid data found in dielectric file: %s\n", file.c_str());
        return false;
    }
    if (best > 0.1 * f)
        fprintf(stderr, "Warning: No exact frequency match in %s. Using closest frequency (diff: %.2e Hz)\n",
                file.c_str(), best);
    return true;
}

// Material files and switchxz (fdtd.c:2248-2470).
void load_materials(Cfg &c) {
    for (size_t i = 0; i < c.mats.size(); ++i) {
        Mat &m = c.mats[i];
        const int n = m.natom;
        if (n > 0 && !m.born_file.empty()) {
            printf("read BORN charges\n");
            double re[9];
            if (!read_outcar_eps(m.born_file, n, re)) die("Failed to read Born charges from %s", m.born_file.c_str());
            if (m.born_eps) {  // born= overwrites eps and zeroes eps_im
                memcpy(m.re, re, sizeof(re));
                memset(m.im, 0, sizeof(m.im));
                if (m.switchxz == 1) {
                    switch_xz(m.re, 3);
                    switch_xz(m.im, 3);
                }
            }
        }
        if (n > 0) {
            if (!m.qezv_set) die("Error: QeZV is not set!");
            if (!m.qezm_set) die("Error: QeZM is not set!");
            if (!m.dmm_set) die("Error: DMM is not set!");
            printf("read QeZV\n");
            m.qezv = read_numbers(m.qezv_file, 9 * n);
            printf("Read QeZM\n");
            m.qezm = read_numbers(m.qezm_file, 9 * n);
            printf("Read DMM\n");
            m.dmm = read_numbers(m.dmm_file, 9 * n * n);
            if (m.switchxz == 1) {
                for (int a = 0; a < n; ++a) {
                    switch_xz(&m.qezv[9 * a], 3);
                    switch_xz(&m.qezm[9 * a], 3);
                }
                for (int a = 0; a < n; ++a)
                    for (int b = 0; b < n; ++b) switch_xz(&m.dmm[(size_t)a * 9 * n + b * 3], 3 * n);
            }
        }
    }
    for (size_t i = 0; i < c.mats.size(); ++i) {
        Mat &m = c.mats[i];
        if (m.use_diel) {  // a dielectric table overrides keys and OUTCAR
            if (!read_dielectric(m.diel_file, c.fsrc, m.re, m.im))
                die("Failed to read dielectric function from %s", m.diel_file.c_str());
            if (m.switchxz == 1) {
                switch_xz(m.re, 3);
                switch_xz(m.im, 3);
            }
            printf("Loaded dielectric function for material %zu from %s at frequency %.2e Hz\n", i,
                   m.diel_file.c_str(), c.fsrc);
        }
        if (m.natom == 0 && !m.use_diel && m.switchxz != 0)
            die("Error: Please switch X and Z by hand for this simple material");
    }
}

int component_of(const std::string &s) {
    static const char *names[] = {"Ex", "Ey", "Ez", "Dx", "Dy", "Dz"};
    for (int q = 0; q < 6; ++q)
        if (s == names[q]) return q;
    return s == "E" ? PHPFDTD_EMAG : PHPFDTD_EX;
}

void check(int rc, const char *what) {
    if (rc < 0) {
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, phpfdtd_last_error());
        exit(1);
    }
}

double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

}  // namespace

int main(int argc, char **argv) {
    const char *cfg_path = nullptr;
    int precision = PHPFDTD_FP64;
    bool precise = false;
    for (int a = 1; a < argc; ++a) {
        if (!strcmp(argv[a], "--fp32")) precision = PHPFDTD_FP32;
        else if (!strcmp(argv[a], "--fp32-rv64")) precision = PHPFDTD_FP32_RV64;
        else if (!strcmp(argv[a], "--precise")) precise = true;
        else cfg_path = argv[a];
    }
    if (!cfg_path) {
        fprintf(stderr, "usage: %s case.cfg [--fp32 | --fp32-rv64] [--precise]\n", argv[0]);
        return 2;
    }
    Cfg c = read_config(cfg_path);
    load_materials(c);
    printf("Grid %d x %d x %d, %d steps, dt = %.6e s, %s form\n", c.nx, c.ny, c.nz, c.tsteps, c.dt,
           precision == PHPFDTD_FP32 ? "FP32 precision" : precision == PHPFDTD_FP32_RV64 ? "FP32 precision (r, v FP64)"
                                                                                    : "FP64 plain");

    // medium_map.bin: int32 nx, ny, nz, then ids with z fastest (fdtd.c:718-959)
    std::vector<int32_t> med((size_t)c.nx * c.ny * c.nz);
    {
        printf("Using medium file: %s\n", c.medium_file.c_str());
        FILE *f = fopen(c.medium_file.c_str(), "rb");
        if (!f) die("Error opening medium file %s", c.medium_file.c_str());
        int32_t dims[3];
        if (fread(dims, 4, 3, f) != 3) die("Error reading medium file header");
        if (dims[0] != c.nx || dims[1] != c.ny || dims[2] != c.nz) {
            fprintf(stderr, "Medium file dimensions (%d,%d,%d) do not match simulation (%d,%d,%d)\n", dims[0],
                    dims[1], dims[2], c.nx, c.ny, c.nz);
            return 1;
        }
        if (fread(med.data(), 4, med.size(), f) != med.size()) die("Error reading medium data");
        fclose(f);
    }

    phpfdtd_grid g;
    memset(&g, 0, sizeof(g));
    g.nx = c.nx; g.ny = c.ny; g.nz = c.nz; g.tsteps = c.tsteps;
    g.dx = c.dx; g.dy = c.dy; g.dz = c.dz; g.dt = c.dt;
    g.eps0 = c.eps0; g.mu0 = c.mu0;
    g.pml_thickness = c.pml; g.pml_reflection = c.pml_r;
    g.pml_order = 3.0; g.kappa_max = 7.0; g.alpha_max = 0.05;  // init_pml (fdtd.c:975-977)
    g.f_source = c.fsrc;
    g.precision = precision;
    std::vector<phpfdtd_material> pm(c.mats.size());
    for (size_t i = 0; i < c.mats.size(); ++i) {
        const Mat &m = c.mats[i];
        memset(&pm[i], 0, sizeof(pm[i]));
        memcpy(pm[i].eps_re, m.re, sizeof(m.re));
        memcpy(pm[i].eps_im, m.im, sizeof(m.im));
        pm[i].sigma = m.sigma;
        pm[i].anglez = m.anglez;
        pm[i].anglex = m.anglex;
        pm[i].use_ade = m.use_ade;
        pm[i].natom = m.natom;
        pm[i].qezv = m.qezv.empty() ? nullptr : m.qezv.data();
        pm[i].qezm = m.qezm.empty() ? nullptr : m.qezm.data();
        pm[i].dmm = m.dmm.empty() ? nullptr : m.dmm.data();
    }
    phpfdtd_ctx *ctx = nullptr;
    check(phpfdtd_create(&g, pm.data(), (int)pm.size(), med.data(), &ctx), "phpfdtd_create");
    if (!c.src.empty()) check(phpfdtd_set_sources(ctx, (int)c.src.size(), c.src.data()), "phpfdtd_set_sources");

    const bool mon_in = c.mon && c.mx >= 0 && c.mx < c.nx && c.my >= 0 && c.my < c.ny && c.mz >= 0 && c.mz < c.nz;
    int mon_natom = 0;
    if (mon_in) {
        const int id = med[((size_t)c.mx * c.ny + c.my) * c.nz + c.mz];
        mon_natom = c.mats[id].natom;
        check(phpfdtd_set_monitor(ctx, c.mx, c.my, c.mz, c.matom, c.mdir, 1), "phpfdtd_set_monitor");
    }
    const bool mon_rv = c.matom < mon_natom;
    const char *fmt = precise ? "%.17e" : "%e";
    FILE *mon_f = nullptr, *full_f = nullptr;
    if (mon_in) {
        mon_f = fopen(c.mon_file.c_str(), "w");  // an empty name opens nothing
        if (mon_f)
            fprintf(mon_f, mon_rv ? "# Time(s)\tEx\tEy\tEz\tDx\tDy\tDz\tHx\tHy\tHz\tR\tV\n"
                                  : "# Time(s)\tEx\tEy\tEz\tDx\tDy\tDz\tHx\tHy\tHz\n");
    }
    if (c.full && mon_in) {
        full_f = fopen(c.full_file.c_str(), "w");
        if (full_f) {
            fprintf(full_f, "# Time(s)\tEx\tEy\tEz\tDx\tDy\tDz\tHx\tHy\tHz");
            for (int a = 0; a < mon_natom; ++a)
                fprintf(full_f, "\tr%d_x\tr%d_y\tr%d_z\tv%d_x\tv%d_y\tv%d_z", a, a, a, a, a, a);
            fprintf(full_f, "\n");
        }
    }
    FILE *src_f = fopen("source.txt", "w");
    if (!src_f) fprintf(stderr, "Error opening source.txt for writing\n");
    const int comp = component_of(c.out_comp);
    const size_t plane = c.axis == 'z' ? (size_t)c.nx * c.ny : c.axis == 'x' ? (size_t)c.ny * c.nz : (size_t)c.nx * c.nz;
    const int d0 = c.axis == 'x' ? c.ny : c.nx, d1 = c.axis == 'z' ? c.ny : c.nz;
    std::vector<float> slice(plane);
    std::vector<double> rows((size_t)(c.out_int > 0 ? c.out_int : 1) * 12 + 12);

    auto flush_monitor = [&]() {
        int got = 0;
        do {
            check(phpfdtd_read_monitor(ctx, rows.data(), (int)(rows.size() / 12), &got), "phpfdtd_read_monitor");
            if (!mon_f) continue;
            for (int r = 0; r < got; ++r) {
                const double *row = &rows[(size_t)r * 12];
                const int ncol = mon_rv ? 12 : 10;
                for (int q = 0; q < ncol; ++q) {
                    fprintf(mon_f, fmt, row[q]);
                    fputc(q + 1 < ncol ? '\t' : '\n', mon_f);
                }
            }
        } while (got > 0);
    };

    printf("Entering iteration:\n");
    fflush(stdout);
    const double t_start = now_s();
    for (int t = 0; t < c.tsteps;) {
        const bool out = c.out_int > 0 && t % c.out_int == 0;
        if (out && src_f) {  // source.txt line (fdtd.c:3359-3367)
            fprintf(src_f, "%.4e  ", t * c.dt);
            for (const auto &s : c.src) fprintf(src_f, "%.4e  ", phpfdtd_source_value(&s, t * c.dt));
            fprintf(src_f, "\n");
            fflush(src_f);
        }
        if (!out) {  // run up to the next output step in one enqueue
            int n = c.out_int > 0 ? c.out_int - t % c.out_int : c.tsteps - t;
            if (t + n > c.tsteps) n = c.tsteps - t;
            check(phpfdtd_run(ctx, n), "phpfdtd_run");
            t += n;
            flush_monitor();
            continue;
        }
        check(phpfdtd_run(ctx, 1), "phpfdtd_run");
        phpfdtd_diag d;
        const int rc = phpfdtd_sync(ctx, &d);
        if (rc < 0 && rc != PHPFDTD_EDIVERGED) check(rc, "phpfdtd_sync");
        printf("it %d/%d (%.1f%%), t = %.5e s,|E|_max = %.2e V/m,", t, c.tsteps, 100.0 * t / c.tsteps, t * c.dt,
               d.max_e);
        printf("r_max = %.2e A,", d.max_r / 1e-10);
        if (t > 0) {
            const double el = now_s() - t_start, eta = el / t * (c.tsteps - t);
            const int eh = (int)(el / 3600), em = (int)((el - eh * 3600) / 60), es = (int)(el - eh * 3600 - em * 60);
            const int rh = (int)(eta / 3600), rm = (int)((eta - rh * 3600) / 60), rs = (int)(eta - rh * 3600 - rm * 60);
            printf(" elapsed: %03d:%02d:%02d, ETA: %03d:%02d:%02d", eh, em, es, rh, rm, rs);
        }
        printf("\n");
        fflush(stdout);
        if (d.diverged) {  // fdtd.c:3626-3667
            printf("ERROR: r field too large at step %d\n", t);
            printf("Max r value: %e\n", d.max_r);
            printf("Location: i=%d, j=%d, k=%d (global k=%d)\n", d.r_i, d.r_j, d.r_k, d.r_k);
            printf("Atom: %d, Direction: %d\n", d.r_atom, d.r_dir);
            double r[96], v[96];
            const int na = phpfdtd_read_ions(ctx, d.r_i, d.r_j, d.r_k, r, v);
            const int q = d.r_atom * 3 + d.r_dir;
            if (na > 0) {
                printf("v value: %e\n", v[q]);
                // a = QeZM^T E - DMM r at the cell itself: the reference indexes its
                // halo-offset field pointers with the loop's k here (fdtd.c:3641-3645),
                // i.e. physical plane k, unlike the monitor's k + halo (fdtd.c:3814)
                double f9[9];
                check(phpfdtd_read_point(ctx, d.r_i, d.r_j, d.r_k, f9), "phpfdtd_read_point");
                const Mat &m = c.mats[med[((size_t)d.r_i * c.ny + d.r_j) * c.nz + d.r_k]];
                double a = 0;
                for (int e = 0; e < 3; ++e) a += m.qezm[d.r_atom * 9 + e * 3 + d.r_dir] * f9[e];
                for (int j = 0; j < 3 * na; ++j) a -= m.dmm[(size_t)q * 3 * na + j] * r[j];
                printf("a value: %e\n", a);
            }
            phpfdtd_destroy(ctx);
            return 1;
        }
        flush_monitor();
        // slice (fdtd.c:3670-3726, 999-1400)
        char name[128];
        snprintf(name, sizeof(name), "field_slice_%c%d_%08d.bin", c.axis, c.pos, t);
        if (phpfdtd_read_slice(ctx, comp, c.axis, c.pos, 1, slice.data()) == 0) {
            FILE *f = fopen(name, "wb");
            if (f) {
                const int32_t dims[2] = {d0, d1};
                fwrite(dims, 4, 2, f);
                fwrite(slice.data(), 4, plane, f);
                fclose(f);
            } else {
                perror("Error opening field file");
            }
        } else {
            fprintf(stderr, "%s\n", phpfdtd_last_error());
        }
        // full monitor (fdtd.c:3728-3776): fields at plane z+1, r and v at z
        if (full_f) {
            double f9[9], r[96], v[96];
            check(phpfdtd_read_point(ctx, c.mx, c.my, c.mz + 1, f9), "phpfdtd_read_point");
            const int na = mon_natom > 0 ? phpfdtd_read_ions(ctx, c.mx, c.my, c.mz, r, v) : 0;
            fprintf(full_f, fmt, t * c.dt);
            for (int q = 0; q < 9; ++q) {
                fputc('\t', full_f);
                fprintf(full_f, fmt, f9[q]);
            }
            for (int a = 0; a < na; ++a) {
                for (int q = 0; q < 3; ++q) {
                    fputc('\t', full_f);
                    fprintf(full_f, fmt, r[3 * a + q]);
                }
                for (int q = 0; q < 3; ++q) {
                    fputc('\t', full_f);
                    fprintf(full_f, fmt, v[3 * a + q]);
                }
            }
            fputc('\n', full_f);
        }
        t += 1;
    }
    check(phpfdtd_sync(ctx, nullptr), "phpfdtd_sync");
    flush_monitor();
    const double t_loop = now_s() - t_start;
    if (src_f) fclose(src_f);
    if (mon_f) {
        fclose(mon_f);
        printf("Monitor data saved to %s\n", c.mon_file.c_str());
    }
    if (full_f) {
        fclose(full_f);
        printf("Full monitor data saved to %s\n", c.full_file.c_str());
    }
    printf("\nSimulation completed in %.2f seconds\n", t_loop);
    phpfdtd_destroy(ctx);
    return 0;
}
