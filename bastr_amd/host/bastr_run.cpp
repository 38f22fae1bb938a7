// bastr_run -- drop-in host driver for the MI355X step library.
//
// Mirrors the reference program (src/Bastr.F90) and its host-side helpers
// with the same file formats, so an existing Bastr case directory runs
// unchanged:
//   datin/input.dat                 readinput            (readwrite.F90:19-112)
//   datin/flowini2d.h5 / 3d.h5      read_initial_field   (readwrite.F90:114-181)
//   outdat/flowfieldNNNN.h5         writeflfed           (readwrite.F90:222-257)
//                                   read_continue_field  (readwrite.F90:183-220)
//   log/stat?d_*.dat, factor_output listinit / listwrite (utility.F90:23-164)
//   bastr_run [NNNN]                restart switch       (Bastr.F90:44-59)
// The time loop (mainloop, solution.F90:20-170) runs on the GPU through the
// C-ABI of include/bastr.h; this program only reads, writes and schedules.
//
// Post-processing commands of the reference's Bastrpp program (Bastrpp.F90):
//   bastr_run hitgen [seed [k0 [energy]]]   hitgen2d/3d (:601-808) with a
//       reproducible seed: datin/flowini{2,3}d.h5 for the grid of
//       datin/input.dat; k0 and energy default to forcek and targetenergy
//   bastr_run spectra2d NNNN    spectra2d (:307-416): pp/SpectraNNNN.dat and
//       pp/Spectra_auxNNNN.dat from outdat/flowfieldNNNN.h5
//   bastr_run transfer2d NNNN   transfer2d (:174-305): pp/TransferNNNN.dat
//   bastr_run transfer3d NNNN   transfer3d (:418-599): pp/TransferNNNN.dat
#include <dlfcn.h>
#include <hdf5.h>  // types and prototypes only: the library is opened with dlopen below
#include <sys/stat.h>

#include <cctype>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/bastr.h"

namespace {

[[noreturn]] void stop_msg(const std::string& msg, int code) {
    std::printf("STOP %s\n", msg.c_str());
    std::fflush(stdout);
    std::exit(code);
}

// The image's HDF5 (a C library, with its own libz) is opened by absolute path
// so that its directory never enters this program's library search path
// (which would shadow the C++ runtime the ROCm libraries need).
struct H5Api {
    decltype(&::H5open) open;
    decltype(&::H5Eset_auto2) eset_auto2;
    decltype(&::H5Fcreate) fcreate;
    decltype(&::H5Fopen) fopen;
    decltype(&::H5Fclose) fclose;
    decltype(&::H5Dopen2) dopen2;
    decltype(&::H5Dcreate2) dcreate2;
    decltype(&::H5Dread) dread;
    decltype(&::H5Dwrite) dwrite;
    decltype(&::H5Dclose) dclose;
    decltype(&::H5Dget_space) dget_space;
    decltype(&::H5Screate_simple) screate_simple;
    decltype(&::H5Sclose) sclose;
    decltype(&::H5Sselect_hyperslab) sselect_hyperslab;
    decltype(&::H5Sget_simple_extent_ndims) sget_ndims;
    hid_t native_double = -1, native_int = -1;
};
H5Api H5;
const unsigned kAccTrunc = 0x0002u, kAccRdonly = 0x0000u;  // H5F_ACC_TRUNC, H5F_ACC_RDONLY

void h5_load() {
    const char* env = std::getenv("BASTR_HDF5_LIB");
    const char* path = env ? env : "/opt/conda/lib/libhdf5.so.103";
    void* lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!lib) stop_msg(std::string("cannot load HDF5 (") + path + "): " + dlerror(), 2);
    auto sym = [&](const char* n) {
        void* f = dlsym(lib, n);
        if (!f) stop_msg(std::string("HDF5 symbol missing: ") + n, 2);
        return f;
    };
    H5.open = (decltype(H5.open))sym("H5open");
    H5.eset_auto2 = (decltype(H5.eset_auto2))sym("H5Eset_auto2");
    H5.fcreate = (decltype(H5.fcreate))sym("H5Fcreate");
    H5.fopen = (decltype(H5.fopen))sym("H5Fopen");
    H5.fclose = (decltype(H5.fclose))sym("H5Fclose");
    H5.dopen2 = (decltype(H5.dopen2))sym("H5Dopen2");
    H5.dcreate2 = (decltype(H5.dcreate2))sym("H5Dcreate2");
    H5.dread = (decltype(H5.dread))sym("H5Dread");
    H5.dwrite = (decltype(H5.dwrite))sym("H5Dwrite");
    H5.dclose = (decltype(H5.dclose))sym("H5Dclose");
    H5.dget_space = (decltype(H5.dget_space))sym("H5Dget_space");
    H5.screate_simple = (decltype(H5.screate_simple))sym("H5Screate_simple");
    H5.sclose = (decltype(H5.sclose))sym("H5Sclose");
    H5.sselect_hyperslab = (decltype(H5.sselect_hyperslab))sym("H5Sselect_hyperslab");
    H5.sget_ndims = (decltype(H5.sget_ndims))sym("H5Sget_simple_extent_ndims");
    if (H5.open() < 0) stop_msg("H5open failed", 2);
    H5.native_double = *(hid_t*)sym("H5T_NATIVE_DOUBLE_g");
    H5.native_int = *(hid_t*)sym("H5T_NATIVE_INT_g");
    H5.eset_auto2(H5E_DEFAULT, nullptr, nullptr);
}

// ------------------------------------------------------------------ input.dat
struct Input {
    int ia = 0, ja = 0, ka = 0;
    double ref_tem = 0, reynolds = 0;
    int maxstep = 0;
    double deltat = 0;
    bool lwsequ = false;
    int feqwsequ = 1;
    bool lwspectra = false;
    int feqwspe = 1, kmax = 0, timemethod = 1;
    int forcemethod = 0, forcek = 0;
    double target_energy = 0;
    bool lprojectd = false;
    int initialmethod = 0;
};

// list-directed item splitting: commas and blanks separate values
std::vector<std::string> items(const std::string& line) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : line) {
        if (c == ',' || std::isspace((unsigned char)c)) {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

double fdouble(std::string s) {
    for (char& c : s)
        if (c == 'd' || c == 'D') c = 'e';
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) stop_msg("readinput: bad real '" + s + "'", 2);
    return v;
}

int fint(const std::string& s) {
    char* end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str()) stop_msg("readinput: bad integer '" + s + "'", 2);
    return (int)v;
}

bool flogical(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    if (!s.empty() && s[0] == '.') s = s.substr(1);
    if (!s.empty() && s[0] == 't') return true;
    if (!s.empty() && s[0] == 'f') return false;
    stop_msg("readinput: bad logical '" + s + "'", 2);
}

// read(fh,'(////)') skips 5 records, read(fh,'(/)') skips 2 (readwrite.F90:35-44)
Input readinput(const std::string& path) {
    std::ifstream f(path);
    if (!f) stop_msg("cannot open " + path, 2);
    std::vector<std::string> lines;
    std::string l;
    while (std::getline(f, l)) lines.push_back(l);
    size_t p = 0;
    auto skip = [&](int n) { p += n; };
    auto next = [&](size_t nmin) {
        // list-directed reads continue on the following records if needed
        std::vector<std::string> v;
        while (v.size() < nmin) {
            if (p >= lines.size()) stop_msg("readinput: end of file in " + path, 2);
            auto t = items(lines[p++]);
            v.insert(v.end(), t.begin(), t.end());
        }
        return v;
    };
    Input in;
    skip(5);
    auto a = next(3);
    in.ia = fint(a[0]); in.ja = fint(a[1]); in.ka = fint(a[2]);
    skip(2);
    a = next(2);
    in.ref_tem = fdouble(a[0]); in.reynolds = fdouble(a[1]);
    skip(2);
    a = next(8);
    in.maxstep = fint(a[0]); in.deltat = fdouble(a[1]); in.lwsequ = flogical(a[2]); in.feqwsequ = fint(a[3]);
    in.lwspectra = flogical(a[4]); in.feqwspe = fint(a[5]); in.kmax = fint(a[6]); in.timemethod = fint(a[7]);
    skip(2);
    a = next(4);
    in.forcemethod = fint(a[0]); in.forcek = fint(a[1]); in.target_energy = fdouble(a[2]);
    in.lprojectd = flogical(a[3]);
    skip(2);
    a = next(1);
    in.initialmethod = fint(a[0]);
    std::printf(" << %s ... done\n", path.c_str());
    return in;
}

// ----------------------------------------------------------------------- logs
// Fortran Ew.d with e=2 exponent digits: [sign]0.<d digits>E<sign><2 digits>
std::string fmt_e(double v, int w = 20, int d = 13) {
    std::string s;
    if (std::isnan(v)) s = "NaN";
    else if (std::isinf(v)) s = v > 0 ? "Infinity" : "-Infinity";
    else {
        const bool neg = std::signbit(v);
        int ex = 0;
        std::string dig;
        if (v == 0) {
            dig.assign(d, '0');
            ex = 0;
        } else {
            char buf[64];
            std::snprintf(buf, sizeof buf, "%.*e", d - 1, std::fabs(v));
            const char* e = std::strchr(buf, 'e');
            dig = std::string(1, buf[0]) + std::string((const char*)buf + 2, e);
            ex = std::atoi(e + 1) + 1;
        }
        if (std::abs(ex) > 99) return std::string(w, '*');
        char eb[8];
        std::snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
        s = std::string(neg ? "-" : "") + "0." + dig + eb;
    }
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

struct Log {
    FILE* f = nullptr;
};

// listinit (utility.F90:23-75): new file with a '(n(1X,A20))' header, or on
// restart resume the existing file at the first record with ns >= nstep.
Log listinit(const std::string& filename, const std::string& firstline, bool lrestart, int64_t nstep) {
    Log lg;
    std::ifstream old(filename);
    if (lrestart && old) {
        std::vector<std::string> keep;
        std::string l;
        std::getline(old, l);
        keep.push_back(l);
        int64_t ns = 0;
        while (std::getline(old, l)) {
            std::istringstream is(l);
            if (!(is >> ns)) break;
            if (ns >= nstep) break;
            keep.push_back(l);
        }
        old.close();
        lg.f = std::fopen(filename.c_str(), "w");
        if (!lg.f) stop_msg("cannot open " + filename, 2);
        for (auto& k : keep) std::fprintf(lg.f, "%s\n", k.c_str());
        std::printf("   ** resume%sat step: %lld\n", filename.c_str(), (long long)ns);
        return lg;
    }
    lg.f = std::fopen(filename.c_str(), "w");
    if (!lg.f) stop_msg("cannot open " + filename, 2);
    for (auto& name : items(firstline)) {
        std::string n = name.substr(0, 20);
        std::fprintf(lg.f, " %20s", n.c_str());
    }
    std::fprintf(lg.f, "\n");
    return lg;
}

// listwrite (utility.F90:87-164): (1X,I20,n(1X,E20.13E2)) with nstep, time first
void listwrite(Log& lg, int64_t nstep, double time, const std::vector<double>& v) {
    std::fprintf(lg.f, " %20lld", (long long)nstep);
    std::fprintf(lg.f, " %s", fmt_e(time).c_str());
    for (double x : v) std::fprintf(lg.f, " %s", fmt_e(x).c_str());
    std::fprintf(lg.f, "\n");
}

// ----------------------------------------------------------------------- HDF5
// A field dataset has the Fortran extent (ia+1, ja+1, ka+1), i.e. the C extent
// (ka+1, ja+1, ia+1); the solver's (ia, ja, ka|1) block sits at offset 0
// (hdf5io.F90:752-802).  Reads take that block (hdf5io.F90:204-256).
void h5read_field(hid_t file, const char* name, int ia, int ja, int nz, double* buf) {
    hid_t d = H5.dopen2(file, name, H5P_DEFAULT);
    if (d < 0) stop_msg(std::string("h5read: no dataset ") + name, 2);
    hid_t fs = H5.dget_space(d);
    hsize_t cnt[3] = {(hsize_t)nz, (hsize_t)ja, (hsize_t)ia}, off[3] = {0, 0, 0};
    if (H5.sget_ndims(fs) != 3) stop_msg(std::string("h5read: ") + name + " is not rank 3", 2);
    H5.sselect_hyperslab(fs, H5S_SELECT_SET, off, nullptr, cnt, nullptr);
    hid_t ms = H5.screate_simple(3, cnt, nullptr);
    if (H5.dread(d, H5.native_double, ms, fs, H5P_DEFAULT, buf) < 0)
        stop_msg(std::string("h5read: failed to read ") + name, 2);
    H5.sclose(ms);
    H5.sclose(fs);
    H5.dclose(d);
}

void h5write_field(hid_t file, const char* name, int ia, int ja, int ka, int nz, const double* buf) {
    hsize_t dimat[3] = {(hsize_t)ka + 1, (hsize_t)ja + 1, (hsize_t)ia + 1};
    hid_t fs = H5.screate_simple(3, dimat, nullptr);
    hid_t d = H5.dcreate2(file, name, H5.native_double, fs, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hsize_t cnt[3] = {(hsize_t)nz, (hsize_t)ja, (hsize_t)ia}, off[3] = {0, 0, 0};
    H5.sselect_hyperslab(fs, H5S_SELECT_SET, off, nullptr, cnt, nullptr);
    hid_t ms = H5.screate_simple(3, cnt, nullptr);
    if (d < 0 || H5.dwrite(d, H5.native_double, ms, fs, H5P_DEFAULT, buf) < 0)
        stop_msg(std::string("h5write: failed to write ") + name, 2);
    H5.sclose(ms);
    H5.sclose(fs);
    H5.dclose(d);
}

// scalar datasets of dims [1] (h5_write1int / h5_write1rl8, hdf5io.F90:1009-1184)
void h5write_scalar(hid_t file, const char* name, hid_t type, const void* v) {
    hsize_t one = 1;
    hid_t s = H5.screate_simple(1, &one, nullptr);
    hid_t d = H5.dcreate2(file, name, type, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (d < 0 || H5.dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0)
        stop_msg(std::string("h5write: failed to write ") + name, 2);
    H5.dclose(d);
    H5.sclose(s);
}

void h5read_scalar(hid_t file, const char* name, hid_t type, void* v) {
    hid_t d = H5.dopen2(file, name, H5P_DEFAULT);
    if (d < 0 || H5.dread(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0)
        stop_msg(std::string("h5read: no scalar ") + name, 2);
    H5.dclose(d);
}

std::string stepname(int filenumb) {
    char b[32];
    std::snprintf(b, sizeof b, "outdat/flowfield%04d.h5", filenumb);
    return b;
}

// ------------------------------------------------------------------- solver
struct Run {
    Input in;
    int nd = 2, nz = 1;
    size_t n = 0;
    bastr_handle* h = nullptr;
    std::vector<double> u[3];
    int allkmax = 0;
};

void check(int rc, const char* what) {
    if (rc < 0) stop_msg(std::string(what) + ": " + bastr_last_error(), 1);
}

void writeflfed(Run& r, int filenumb, int64_t nstep, double time) {
    check(bastr_get_fields(r.h, r.u[0].data(), r.u[1].data(), r.nd == 3 ? r.u[2].data() : nullptr),
          "bastr_get_fields");
    const std::string fn = stepname(filenumb);
    hid_t f = H5.fcreate(fn.c_str(), kAccTrunc, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) stop_msg(" !! error in h5io_init call h5fcreate_f: " + fn, 2);
    const char* names[3] = {"u1", "u2", "u3"};
    for (int c = 0; c < r.nd; ++c) h5write_field(f, names[c], r.in.ia, r.in.ja, r.in.ka, r.nz, r.u[c].data());
    const int ns = (int)nstep;
    h5write_scalar(f, "nstep", H5.native_int, &ns);
    h5write_scalar(f, "time", H5.native_double, &time);
    H5.fclose(f);
    std::printf(" >> %s ... done\n", fn.c_str());
}

// hitgen: seeded random-phase dilatational field, create_initial_field_2d/3d
// (Bastrpp.F90:601-808) step by step with a SplitMix64 noise source in place of
// the clock-seeded random_number, on the library's fftwlink-compatible
// transforms.
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

int hitgen(const Input& in, double k0, uint64_t seed, double target) {
    const int nd = in.ka == 0 ? 2 : 3;
    const int nz = nd == 3 ? in.ka : 1;
    const size_t n = (size_t)in.ia * in.ja * nz;
    bastr_params p{};
    p.ia = in.ia; p.ja = in.ja; p.ka = in.ka;
    p.reynolds = 1; p.deltat = 1; p.timemethod = 1; p.kmax = 1;
    bastr_handle* h = nullptr;
    check(bastr_create(&p, &h), "bastr_create");
    std::vector<std::complex<double>> z(n);
    for (size_t m = 0; m < n; ++m)
        z[m] = (double)(splitmix64(seed * 0x100000001B3ull + m) >> 11) * 0x1.0p-53;
    check(bastr_fft_forward(h, (double*)z.data()), "bastr_fft_forward");  // phases of the noise
    std::vector<double> theta(n);
    for (size_t m = 0; m < n; ++m) theta[m] = std::arg(z[m]);
    std::vector<double> u[3];
    const char* names[3] = {"u1", "u2", "u3"};
    double energy = 0;
    for (int c = 0; c < nd; ++c) {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < in.ja; ++j)
                for (int i = 0; i < in.ia; ++i) {
                    const size_t m = (size_t)i + (size_t)in.ia * (j + (size_t)in.ja * k);
                    const double k1 = i <= in.ia / 2 ? i : i - in.ia;
                    const double k2 = j <= in.ja / 2 ? j : j - in.ja;
                    const double k3 = nd == 3 ? (k <= in.ka / 2 ? k : k - in.ka) : 0.0;
                    const double kk = std::sqrt(k1 * k1 + k2 * k2 + k3 * k3);
                    if (kk == 0) { z[m] = 0; continue; }
                    const double q = kk / k0, kk2 = kk * kk;
                    const double E = kk >= k0 ? kk2 * kk2 * std::exp(-2.0 * (q * q))
                                              : k0 * k0 * k0 * k0 * k0 * k0 / kk2 * std::exp(-2.0);
                    const double A = nd == 2 ? std::sqrt(E / (2 * M_PI * kk)) : std::sqrt(E / (4 * M_PI * kk * kk));
                    const double kc = c == 0 ? k1 : (c == 1 ? k2 : k3);
                    z[m] = A * kc / kk * std::complex<double>(std::sin(theta[m]), std::cos(theta[m]));
                }
        check(bastr_fft_backward(h, (double*)z.data()), "bastr_fft_backward");
        u[c].resize(n);
        for (size_t m = 0; m < n; ++m) u[c][m] = z[m].real();
    }
    for (size_t m = 0; m < n; ++m)
        for (int c = 0; c < nd; ++c) energy += u[c][m] * u[c][m];
    energy /= (double)n;
    const double f = std::sqrt(target / energy);
    const std::string fn = nd == 2 ? "datin/flowini2d.h5" : "datin/flowini3d.h5";
    hid_t file = H5.fcreate(fn.c_str(), kAccTrunc, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) stop_msg("cannot create " + fn, 2);
    for (int c = 0; c < nd; ++c) {
        for (auto& x : u[c]) x *= f;
        h5write_field(file, names[c], in.ia, in.ja, in.ka, nz, u[c].data());
    }
    H5.fclose(file);
    bastr_destroy(h);
    std::printf(" >> %s ... done\n", fn.c_str());
    return 0;
}

// spectra2d / transfer2d / transfer3d of one dump (Bastrpp.F90:174-599): read
// outdat/flowfieldNNNN.h5, evaluate on the device, write pp/*.dat with the
// dump's nstep and time in the leading columns, as listwrite does.
int postprocess(const Input& in, const std::string& cmd, int filenumb) {
    const int nd = in.ka == 0 ? 2 : 3;
    if ((cmd == "transfer3d") != (nd == 3))
        stop_msg(cmd + " needs a " + (nd == 3 ? "2-D" : "3-D") + " case in datin/input.dat", 1);
    const int nz = nd == 3 ? in.ka : 1;
    const size_t n = (size_t)in.ia * in.ja * nz;
    std::vector<double> u[3];
    const char* names[3] = {"u1", "u2", "u3"};
    const std::string fn = stepname(filenumb);
    hid_t f = H5.fopen(fn.c_str(), kAccRdonly, H5P_DEFAULT);
    if (f < 0) stop_msg(" !! error in h5io_init call h5fopen_f: " + fn, 2);
    for (int c = 0; c < nd; ++c) {
        u[c].assign(n, 0.0);
        h5read_field(f, names[c], in.ia, in.ja, nz, u[c].data());
    }
    int ns = 0;
    double time = 0;
    h5read_scalar(f, "time", H5.native_double, &time);
    h5read_scalar(f, "nstep", H5.native_int, &ns);
    H5.fclose(f);
    std::printf(" >> %s ... done\n", fn.c_str());
    bastr_params p{};
    p.ia = in.ia; p.ja = in.ja; p.ka = in.ka;
    p.reynolds = in.reynolds;
    p.deltat = in.deltat;
    p.timemethod = 1;
    p.kmax = in.kmax;
    const char* prec = std::getenv("BASTR_PRECISION");
    p.precision = (prec && std::string(prec) == "fp32") ? BASTR_PREC_F32 : BASTR_PREC_F64;
    bastr_handle* h = nullptr;
    check(bastr_create(&p, &h), "bastr_create");
    const int K = bastr_allkmax(h);
    mkdir("pp", 0755);
    char num[8];
    std::snprintf(num, sizeof num, "%04d", filenumb);
    const double* u3 = nd == 3 ? u[2].data() : nullptr;
    if (cmd == "spectra2d") {
        std::vector<double> shells(2 * (size_t)K), tot(4);
        check(bastr_pp_spectra(h, u[0].data(), u[1].data(), nullptr, shells.data(), tot.data()), "bastr_pp_spectra");
        std::printf("Calculation finish!\n");
        const std::string out = std::string("pp/Spectra") + num + ".dat";
        Log lg = listinit(out, "ns ti k Es Ed", false, 0);
        for (int b = 0; b < K; ++b) listwrite(lg, ns, time, {(double)(b + 1), shells[b], shells[K + b]});
        std::fclose(lg.f);
        std::printf(" << %s ... done\n", out.c_str());
        const std::string aux = std::string("pp/Spectra_aux") + num + ".dat";
        Log la = listinit(aux, "ns ti Esspe Edspe k2Edspe kM1Edspe", false, 0);
        listwrite(la, ns, time, {tot[0], tot[1], tot[2], tot[3]});
        std::fclose(la.f);
        std::printf(" << %s ... done\n", aux.c_str());
    } else {
        std::vector<double> t(3 * (size_t)K);
        check(bastr_pp_transfer(h, u[0].data(), u[1].data(), u3, t.data()), "bastr_pp_transfer");
        std::printf("Calculation finish!\n");
        const std::string out = std::string("pp/Transfer") + num + ".dat";
        Log lg = listinit(out, "ns ti k T T1 T2", false, 0);
        for (int b = 0; b < K; ++b) listwrite(lg, ns, time, {(double)(b + 1), t[b], t[K + b], t[2 * K + b]});
        std::fclose(lg.f);
        std::printf(" << %s ... done\n", out.c_str());
    }
    bastr_destroy(h);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::string(argv[1]) == "--parse") {
        // print the parsed datin/input.dat and stop (no device use)
        Input in = readinput("datin/input.dat");
        std::printf("ia=%d ja=%d ka=%d ref_tem=%.17g reynolds=%.17g maxstep=%d deltat=%.17g lwsequ=%d "
                    "feqwsequ=%d lwspectra=%d feqwspe=%d kmax=%d timemethod=%d forcemethod=%d forcek=%d "
                    "target_energy=%.17g lprojectd=%d initialmethod=%d\n",
                    in.ia, in.ja, in.ka, in.ref_tem, in.reynolds, in.maxstep, in.deltat, (int)in.lwsequ,
                    in.feqwsequ, (int)in.lwspectra, in.feqwspe, in.kmax, in.timemethod, in.forcemethod,
                    in.forcek, in.target_energy, (int)in.lprojectd, in.initialmethod);
        return 0;
    }
    h5_load();
    Input in = readinput("datin/input.dat");
    std::printf(" ia: %d ja: %d ka: %d\n", in.ia, in.ja, in.ka);
    if (argc >= 2 && std::string(argv[1]) == "hitgen") {
        const uint64_t seed = argc >= 3 ? std::strtoull(argv[2], nullptr, 10) : 1;
        const double k0 = argc >= 4 ? std::atof(argv[3]) : (double)in.forcek;
        const double energy = argc >= 5 ? std::atof(argv[4]) : in.target_energy;
        if (!(k0 > 0) || !(energy > 0)) stop_msg("hitgen: k0 and energy must be > 0", 2);
        return hitgen(in, k0, seed, energy);
    }
    if (argc >= 2 && (std::string(argv[1]) == "spectra2d" || std::string(argv[1]) == "transfer2d" ||
                      std::string(argv[1]) == "transfer3d")) {
        std::printf(" ** pp command: %s\n", argv[1]);
        if (argc < 3) stop_msg(std::string(argv[1]) + ": file number missing", 2);
        return postprocess(in, argv[1], std::atoi(argv[2]));
    }
    if (in.ka == 0 && in.ja == 0) stop_msg("1D case is not implemented yet!", 1);
    if (in.timemethod != 1 && in.timemethod != 2) stop_msg("Undefined time method", 0);
    mkdir("log", 0755);
    mkdir("outdat", 0755);

    Run r;
    r.in = in;
    r.nd = in.ka == 0 ? 2 : 3;
    r.nz = r.nd == 3 ? in.ka : 1;
    r.n = (size_t)in.ia * in.ja * r.nz;
    for (int c = 0; c < r.nd; ++c) r.u[c].assign(r.n, 0.0);
    bastr_params p{};
    p.ia = in.ia; p.ja = in.ja; p.ka = in.ka;
    p.reynolds = in.reynolds;
    p.deltat = in.deltat;
    p.timemethod = in.timemethod;
    p.forcemethod = in.forcemethod;
    p.forcek = in.forcek;
    p.target_energy = in.target_energy;
    p.lproject = in.lprojectd ? 1 : 0;
    p.kmax = in.kmax;
    const char* prec = std::getenv("BASTR_PRECISION");
    p.precision = (prec && std::string(prec) == "fp32") ? BASTR_PREC_F32 : BASTR_PREC_F64;
    check(bastr_create(&p, &r.h), "bastr_create");
    r.allkmax = bastr_allkmax(r.h);
    const double nu = in.reynolds == 0.0 ? 0.0 : 1.0 / in.reynolds;
    std::printf(" nu= %.17g\n allkmax= %d\n", nu, r.allkmax);

    // quantity_prepare (solution.F90:1288-1304)
    int64_t nstep = 0;
    double time = 0.0;
    int maxstep = in.maxstep;
    int nxtwsequ = std::min(in.feqwsequ, maxstep);
    int nxtwspe = 0;
    int filenumb = 1;
    bool lrestart = false;
    const char* names[3] = {"u1", "u2", "u3"};
    if (argc < 2 || std::strlen(argv[1]) == 0) {
        std::printf("  start from zero.... \n");
        if (in.initialmethod == 1) {
            const std::string fn = r.nd == 2 ? "datin/flowini2d.h5" : "datin/flowini3d.h5";
            hid_t f = H5.fopen(fn.c_str(), kAccRdonly, H5P_DEFAULT);
            if (f < 0) stop_msg(" !! error in h5io_init call h5fopen_f: " + fn, 2);
            for (int c = 0; c < r.nd; ++c) h5read_field(f, names[c], in.ia, in.ja, r.nz, r.u[c].data());
            H5.fclose(f);
            std::printf(" << %s ... done\n", fn.c_str());
        } else if (in.initialmethod == 2 || in.initialmethod == 3) {
            if (r.nd != 2) stop_msg("ndims should be 2!", 1);
            const double PI = 3.14159265358979323846;
            for (int j = 1; j <= in.ja; ++j)
                for (int i = 1; i <= in.ia; ++i) {
                    const size_t m = (size_t)(i - 1) + (size_t)in.ia * (j - 1);
                    const double sx = std::sin(2 * PI * i / in.ia), cx = std::cos(2 * PI * i / in.ia);
                    const double sy = std::sin(2 * PI * j / in.ja), cy = std::cos(2 * PI * j / in.ja);
                    if (in.initialmethod == 2) { r.u[0][m] = sx * cy; r.u[1][m] = -cx * sy; }
                    else { r.u[0][m] = sx * sy; r.u[1][m] = -cx * cy; }
                }
        } else if (in.initialmethod != 0) {
            stop_msg("read_initial_field: Not proper initial method", 1);
        }
    } else {
        lrestart = true;
        filenumb = std::atoi(argv[1]);
        std::printf(" continue to run: start from %d\n", filenumb);
        const std::string fn = stepname(filenumb);
        hid_t f = H5.fopen(fn.c_str(), kAccRdonly, H5P_DEFAULT);
        if (f < 0) stop_msg(" !! error in h5io_init call h5fopen_f: " + fn, 2);
        for (int c = 0; c < r.nd; ++c) h5read_field(f, names[c], in.ia, in.ja, r.nz, r.u[c].data());
        int ns = 0;
        h5read_scalar(f, "nstep", H5.native_int, &ns);
        h5read_scalar(f, "time", H5.native_double, &time);
        H5.fclose(f);
        nstep = ns;
        filenumb = filenumb + 1;
        nxtwsequ = (int)std::min<int64_t>(nstep + in.feqwsequ, maxstep);
        nxtwspe = (int)std::min<int64_t>(nstep + in.feqwspe, maxstep);
        std::printf(" >> %s ... done\n", fn.c_str());
    }
    check(bastr_set_fields(r.h, r.u[0].data(), r.u[1].data(), r.nd == 3 ? r.u[2].data() : nullptr), "set_fields");
    check(bastr_set_time(r.h, nstep, time), "set_time");

    const std::string d = r.nd == 2 ? "2d" : "3d";
    Log lf = listinit("log/stat" + d + "_specall.dat", "ns ti Es Ed k2Ed kMEd", lrestart, nstep);
    Log lg = listinit("log/stat" + d + "_spect.dat", "ns ti k Es Ed", lrestart, nstep);
    Log la = listinit("log/stat" + d + "_velgrad.dat",
                      "ns ti umumtheta2 umumijji u2theta dissp etamin Tay ReTay Kol", lrestart, nstep);
    Log lo;
    const int fm = in.forcemethod;
    if (fm > 0) {
        const char* hdr = fm == 1 ? "ns ti factor" : ((fm == 2 || fm == 5) ? "ns ti factor E" : "ns ti factor");
        lo = listinit("log/factor_output.dat", hdr, lrestart, nstep);
    }

    // mainloop (solution.F90:54-97 / 124-169), in chunks that end where the
    // reference writes a field or a spectrum
    const int K = r.allkmax;
    std::vector<bastr_diag> diags;
    std::vector<double> spec;
    int64_t dumped_at = -1;  // nstep of the last scheduled dump
    while (nstep <= maxstep) {
        if (in.lwsequ && nstep == nxtwsequ) {
            // the reference dumps inside iteration nstep, after its diagnostics
            // and before its step: the device field is exactly that field
            writeflfed(r, filenumb, nstep, time);
            dumped_at = nstep;
            filenumb = filenumb + 1;
            nxtwsequ = std::min(nxtwsequ + in.feqwsequ, maxstep);
            std::printf(" Next print is %d\n", nxtwsequ);
        }
        // the chunk ends before the next dump; spectra of every iteration come back anyway
        int64_t stop_at = std::min<int64_t>(maxstep, nstep + 999);
        if (in.lwsequ && nxtwsequ > nstep) stop_at = std::min<int64_t>(stop_at, nxtwsequ - 1);
        const int nit = (int)(stop_at - nstep + 1);
        diags.assign(nit, bastr_diag{});
        spec.assign((size_t)nit * 3 * K, 0.0);
        const int rc = bastr_step(r.h, nit, diags.data(), spec.data());
        check(rc, "bastr_step");
        int done = nit;
        if (rc == BASTR_STOP_NAN)
            for (int s = 0; s < nit; ++s)
                if (diags[s].nan_flag) { done = s + 1; break; }
        for (int s = 0; s < done; ++s) {
            const bastr_diag& g = diags[s];
            listwrite(lf, g.nstep, g.time, {g.Esspe, g.Edspe, g.k2Edspe, g.kMEdspe});
            const bool nan = g.nan_flag != 0;
            // 3-D also writes the spectrum of a NaN iteration (solution.F90:1095)
            if ((in.lwspectra && g.nstep == nxtwspe) || (r.nd == 3 && nan)) {
                const double* sp = spec.data() + (size_t)s * 3 * K;
                for (int b = 0; b < K; ++b) listwrite(lg, g.nstep, g.time, {sp[b], sp[K + b], sp[2 * K + b]});
                nxtwspe = std::min(nxtwspe + in.feqwspe, maxstep);
            }
            listwrite(la, g.nstep, g.time,
                      {g.umumtheta2, g.umumijji, g.u2theta, g.dissp, g.eta_min, g.Taylength, g.ReTay, g.Kollength});
            std::printf(" nstep, Es, Ed, Eall= %20lld %24.16E %24.16E %24.16E\n", (long long)g.nstep, g.Esspe,
                        g.Edspe, g.Esspe + g.Edspe);
            if (nan) {
                // writeflfed then stop (solution.F90:69-81); one condition covers
                // both the scheduled dump and the NaN, so an iteration that was
                // just dumped is not written a second time
                if (g.nstep != dumped_at) writeflfed(r, filenumb, g.nstep, g.time);
                std::fflush(nullptr);
                stop_msg("\"E\" is a NaN", 0);
            }
            if (fm == 1) listwrite(lo, g.nstep + 1, g.time + in.deltat, {g.force_factor});
            else if (fm == 2 || fm == 5) listwrite(lo, g.nstep + 1, g.time + in.deltat, {g.force_factor, g.force_E});
        }
        check(bastr_get_time(r.h, &nstep, &time), "bastr_get_time");
        // forcing lines use the reference's accumulated time: time(n+1) = time(n) + dt
        std::fflush(nullptr);
    }
    for (Log* l : {&lf, &lg, &la, &lo})
        if (l->f) std::fclose(l->f);
    bastr_destroy(r.h);
    std::printf(" ** End!\n");
    return 0;
}
