// gpk_cli.cpp — `gpk <deck.ini>` driver on the GPU facade.
//
// Runs the in-scope models of the reference CLI (demos/gpk.cpp):
// [run] model = grain_growth (run_grain_growth, 195-282),
// [run] model = solidification (run_solidification, 288-391),
// [run] model = recrystallisation (run_recrystallisation, 477-555) and
// [run] model = precipitation (run_precipitation, 703-797) and
// [run] model = precipitate_breakup (run_precipitate_breakup, 1065-1275), with the same
// deck keys, initial conditions, progress tables and snapshot files
// ([output] format = csv | vtk | ppm | all, dir, prefix, write_initial).
// Optional keys: [gpu] precision = strict32 | fp32 | fp64, [gpu] device.
// Other models exit with status 1, as the reference does for unknown models.
#include "gpk_core.hpp"
#include "gpk_thermo.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

using namespace gpk;

namespace {

// INI decks (utils/input_parser.hpp:52-147): "[section]", "key = value",
// '#' starts a comment, typed getters, a missing key without a default throws.
class Deck {
  public:
    explicit Deck(const std::string& path) {
        std::ifstream f(path);
        if (!f) throw std::runtime_error("InputParser: cannot open '" + path + "'");
        std::string section, line;
        while (std::getline(f, line)) {
            const auto hash = line.find('#');
            if (hash != std::string::npos) line.erase(hash);
            line = trim(line);
            if (line.empty()) continue;
            if (line.front() == '[' && line.back() == ']') {
                section = trim(line.substr(1, line.size() - 2));
                continue;
            }
            const auto eq = line.find('=');
            if (eq == std::string::npos) continue;
            kv_[section + "/" + trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
        }
    }
    bool has(const std::string& s, const std::string& k) const { return kv_.count(s + "/" + k) > 0; }
    std::string str(const std::string& s, const std::string& k, const std::string& def = "") const {
        auto it = kv_.find(s + "/" + k);
        if (it == kv_.end()) {
            if (def.empty()) throw missing(s, k);
            return def;
        }
        return it->second;
    }
    int i(const std::string& s, const std::string& k, int def = -999999) const {
        auto it = kv_.find(s + "/" + k);
        if (it == kv_.end()) {
            if (def == -999999) throw missing(s, k);
            return def;
        }
        return std::stoi(it->second);
    }
    double d(const std::string& s, const std::string& k, double def = 1e300) const {
        auto it = kv_.find(s + "/" + k);
        if (it == kv_.end()) {
            if (def > 1e299) throw missing(s, k);
            return def;
        }
        return std::stod(it->second);
    }
    float f(const std::string& s, const std::string& k, float def = 1e30f) const {
        return static_cast<float>(d(s, k, def < 1e29f ? double(def) : 1e300));
    }
    bool b(const std::string& s, const std::string& k, bool def = false) const {
        auto it = kv_.find(s + "/" + k);
        if (it == kv_.end()) return def;
        std::string v = it->second;
        std::transform(v.begin(), v.end(), v.begin(), ::tolower);
        return v == "true" || v == "1" || v == "yes";
    }

  private:
    std::unordered_map<std::string, std::string> kv_;
    static std::string trim(const std::string& s) {
        const auto b = s.find_first_not_of(" \t\r\n");
        if (b == std::string::npos) return "";
        const auto e = s.find_last_not_of(" \t\r\n");
        return s.substr(b, e - b + 1);
    }
    static std::runtime_error missing(const std::string& s, const std::string& k) {
        return std::runtime_error("InputParser: missing required key [" + s + "] " + k);
    }
};

void mkdir_p(const std::string& path) {
    std::string cur;
    for (char ch : path + "/") {
        if ((ch == '/' || ch == '\\') && !cur.empty()) ::mkdir(cur.c_str(), 0755);
        cur += ch;
    }
}

// Snapshot files.  CSV (utils/csv_writer.hpp:20-56) and legacy VTK
// (utils/vtk_writer.hpp:18-67) carry the reference's layout and number
// formats; PPM images follow utils/ppm_writer.hpp's colouring.
class Writer {
  public:
    explicit Writer(const Deck& cfg)
        : format_(cfg.str("output", "format", "vtk")), dir_(cfg.str("output", "dir", "output")),
          prefix_(cfg.str("output", "prefix", "snap")), compact_(cfg.b("output", "compact", false)) {
        mkdir_p(dir_);
    }
    void write(GPKGrid& grid, int snap, double t) const {
        grid.sync_host();
        char tag[64];
        std::snprintf(tag, sizeof(tag), "%04d", snap);
        const std::string base = dir_ + "/" + prefix_ + "_" + tag;
        if (format_ == "csv" || format_ == "all") csv(grid, base + ".csv", t, snap);
        if (format_ == "vtk" || format_ == "all") vtk(grid, base + ".vtk", t, snap);
        if (format_ == "ppm" || format_ == "all") ppm(grid, base);
    }

  private:
    std::string format_, dir_, prefix_;
    bool compact_;

    void csv(const GPKGrid& grid, const std::string& path, double t, int snap) const {
        std::ofstream f(path);
        if (!f) throw std::runtime_error("CSV: cannot open " + path);
        const float dx = grid.dx;
        f << "# GPK Microstructure CSV\n";
        f << "# snap=" << snap << "  time=" << t << "\n";
        f << "# Nz=" << grid.Nz << "  Ny=" << grid.Ny << "  Nx=" << grid.Nx << "\n";
        f << "# dx=" << dx << " [m]\n";
        if (compact_) f << "# compact=true (only interface + changed cells written)\n";
        f << "# cols: x[m] y[m] z[m] grain_id alpha composition temperature orientation\n";
        char row[256];
        for (int z = 0; z < grid.Nz; ++z)
            for (int y = 0; y < grid.Ny; ++y)
                for (int x = 0; x < grid.Nx; ++x) {
                    const int i = grid.idx(z, y, x);
                    if (compact_ && !grid.mask[i]) continue;
                    std::snprintf(row, sizeof(row), "%.6e %.6e %.6e %d %.6f %.6f %.2f %.6f\n", x * dx,
                                  y * dx, z * dx, grid.phi[i], grid.alpha[i], grid.c[i], grid.T[i],
                                  grid.g[i]);
                    f << row;
                }
    }

    void vtk(const GPKGrid& grid, const std::string& path, double t, int snap) const {
        std::ofstream f(path);
        if (!f) throw std::runtime_error("VTK: cannot open " + path);
        const int N = grid.N();
        f << "# vtk DataFile Version 3.0\n";
        f << "GPK Microstructure snap=" << snap << " time=" << t << "\n";
        f << "ASCII\nDATASET STRUCTURED_POINTS\n";
        f << "DIMENSIONS " << grid.Nx << " " << grid.Ny << " " << grid.Nz << "\n";
        f << "ORIGIN 0 0 0\n";
        f << "SPACING " << grid.dx << " " << grid.dx << " " << grid.dx << "\n";
        f << "POINT_DATA " << N << "\n";
        f << "SCALARS grain_id int 1\nLOOKUP_TABLE default\n";
        for (int i = 0; i < N; ++i) f << grid.phi[i] << ((i % 9 == 8) ? '\n' : ' ');
        f << '\n';
        const std::vector<float>* fields[4] = {&grid.alpha, &grid.c, &grid.T, &grid.g};
        const char* names[4] = {"alpha", "composition", "temperature", "orientation"};
        char buf[32];
        for (int k = 0; k < 4; ++k) {
            f << "SCALARS " << names[k] << " float 1\nLOOKUP_TABLE default\n";
            for (int i = 0; i < N; ++i) {
                std::snprintf(buf, sizeof(buf), "%.6g", (*fields[k])[i]);
                f << buf << ((i % 6 == 5) ? '\n' : ' ');
            }
            f << '\n';
        }
    }

    // Mid-plane (z = Nz/2) images, utils/gpk_io.hpp:77-101 and
    // utils/ppm_writer.hpp: grain colours from a multiplicative label hash,
    // alpha-blended with the competitor colour at interface cells; the
    // composition slice through a piecewise-linear purple-teal-yellow map,
    // written only when c varies on the slice.
    static unsigned char unit_to_byte(float v) {
        return static_cast<unsigned char>(std::min(std::max(v, 0.f), 1.f) * 255.f + 0.5f);
    }
    static void label_colour(int32_t label, unsigned char rgb[3]) {
        const uint32_t h = static_cast<uint32_t>(label) * 2654435761u;
        for (int k = 0; k < 3; ++k) rgb[k] = static_cast<unsigned char>(80 + ((h >> (8 * k)) & 0x7F));
    }
    static void write_p6(const std::string& path, int W, int H, const std::vector<unsigned char>& rgb) {
        std::ofstream f(path, std::ios::binary);
        f << "P6\n" << W << " " << H << "\n255\n";
        f.write(reinterpret_cast<const char*>(rgb.data()), std::streamsize(rgb.size()));
    }

    void ppm(const GPKGrid& grid, const std::string& base) const {
        const int W = grid.Nx, H = grid.Ny;
        const size_t plane = size_t(W) * H, off = size_t(grid.Nz / 2) * plane;
        std::vector<unsigned char> rgb(plane * 3);
        for (size_t p = 0; p < plane; ++p) {
            const size_t i = off + p;
            unsigned char* px = &rgb[3 * p];
            label_colour(grid.phi[i], px);
            if (grid.mask[i] && grid.phi_comp[i] != grid.phi[i]) {
                unsigned char other[3];
                label_colour(grid.phi_comp[i], other);
                const float t = std::min(std::max(grid.alpha[i], 0.f), 1.f);
                for (int k = 0; k < 3; ++k)
                    px[k] = unit_to_byte(t * (px[k] / 255.f) + (1.f - t) * (other[k] / 255.f));
            }
        }
        write_p6(base + "_grains.ppm", W, H, rgb);

        const auto c0 = grid.c.begin() + std::ptrdiff_t(off);
        const float lo = *std::min_element(c0, c0 + std::ptrdiff_t(plane));
        const float hi = *std::max_element(c0, c0 + std::ptrdiff_t(plane));
        if (!(hi > lo)) return;
        const float span = hi - lo;
        for (size_t p = 0; p < plane; ++p) {
            const float t = std::min(std::max((grid.c[off + p] - lo) / span, 0.f), 1.f);
            const bool low = t < 0.5f;
            const float u = t - 0.5f;
            rgb[3 * p + 0] = unit_to_byte(low ? 0.27f + 0.1f * t : 0.2f + 1.6f * u);
            rgb[3 * p + 1] = unit_to_byte(low ? 0.4f * t : 0.2f + 1.2f * u);
            rgb[3 * p + 2] = unit_to_byte(low ? 0.5f - t : 0.5f - u);
        }
        write_p6(base + "_comp.ppm", W, H, rgb);
    }
};

int precision_of(const Deck& cfg) {
    const std::string p = cfg.str("gpu", "precision", "strict32");
    if (p == "strict32") return GPK_PREC_STRICT32;
    if (p == "fp32") return GPK_PREC_FP32;
    if (p == "fp64") return GPK_PREC_FP64;
    throw std::runtime_error("[gpu] precision must be strict32, fp32 or fp64");
}

// demos/gpk.cpp:195-282
int run_grain_growth(const Deck& cfg) {
    const int Nz = cfg.i("grid", "Nz"), Ny = cfg.i("grid", "Ny"), Nx = cfg.i("grid", "Nx");
    const float dx = cfg.f("grid", "dx");
    const double t_end = cfg.d("time", "t_end");
    const int n_snaps = cfg.i("time", "n_snaps");
    const float M0 = cfg.f("thermo", "M0"), sigma0 = cfg.f("thermo", "sigma0");
    const float eps4 = cfg.f("thermo", "eps4", 0.f);
    const int n_grains = cfg.i("grain_growth", "n_grains");
    const int seed = cfg.i("grain_growth", "seed", 42);

    std::printf("=== GPK Grain Growth (MI355X) ===\n");
    std::printf("Grid %d×%d×%d, dx=%.1f µm, %d grains\n", Nz, Ny, Nx, dx * 1e6f, n_grains);
    std::shared_ptr<ThermoModel> thermo;
    if (std::fabs(eps4) > 1e-12f) thermo = std::make_shared<AnisotropicGrowthThermo>(M0, sigma0, eps4);
    else thermo = std::make_shared<IsotropicGrowthThermo>(M0, sigma0);
    GPKGrid grid(Nz, Ny, Nx, dx, thermo, 0.f, cfg.f("geometry", "sigma_cells", 2.f), precision_of(cfg),
                 cfg.i("gpu", "device", 0));
    if (cfg.has("grain_growth", "restart_labels")) {
        auto read_exact = [](const std::string& path, char* dst, std::streamsize bytes) {
            std::ifstream s(path, std::ios::binary);
            if (!s) throw std::runtime_error("Cannot open restart field '" + path + "'");
            s.read(dst, bytes);
            if (s.gcount() != bytes) throw std::runtime_error("Restart field has the wrong size: '" + path + "'");
            char extra = 0;
            if (s.read(&extra, 1)) throw std::runtime_error("Restart field contains trailing data: '" + path + "'");
        };
        read_exact(cfg.str("grain_growth", "restart_labels"), reinterpret_cast<char*>(grid.phi.data()),
                   std::streamsize(grid.phi.size() * sizeof(grid.phi[0])));
        const std::string gpath = cfg.str("grain_growth", "restart_orientations", "none");
        if (gpath != "none")
            read_exact(gpath, reinterpret_cast<char*>(grid.g.data()),
                       std::streamsize(grid.g.size() * sizeof(grid.g[0])));
        std::fill(grid.alpha.begin(), grid.alpha.end(), 1.f);
        grid.time = cfg.d("grain_growth", "restart_time", 0.0);
        grid.mark_dirty();
        std::printf("Restarted categorical state at t=%.4f s\n", grid.time);
    } else {
        grid.voronoi_init(n_grains, uint32_t(seed));
    }
    Writer writer(cfg);
    const bool write_initial = cfg.b("output", "write_initial", false);
    const int off = write_initial ? 1 : 0;
    if (write_initial) writer.write(grid, 0, grid.time);
    std::printf("%-12s %-18s %-10s\n", "Time [s]", "R_char [µm]", "Grains");
    const auto t0 = std::chrono::high_resolution_clock::now();
    for (int snap = 0; snap < n_snaps; ++snap) {
        const double t_snap = t_end * (snap + 1.0) / n_snaps;
        grid.run(t_snap, -1.f,
                 [](const GPKGrid& st, int steps) {
                     std::printf("  progress: step %d, t=%.6g s, grains=%d\n", steps, st.time, st.n_grains());
                     std::fflush(stdout);
                 },
                 50);
        if (grid.n_grains() <= 1) break;
        std::printf("%-12.4f %-18.4f %-10d\n", grid.time, grid.mean_grain_radius() * 1e6, grid.n_grains());
        writer.write(grid, snap + off, grid.time);
    }
    const double wall = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    std::printf("\nWall time: %.2f s\n", wall);
    return 0;
}

// demos/gpk.cpp:288-391
int run_solidification(const Deck& cfg) {
    const int Nz = cfg.i("grid", "Nz"), Ny = cfg.i("grid", "Ny"), Nx = cfg.i("grid", "Nx");
    const float dx = cfg.f("grid", "dx");
    const double t_end = cfg.d("time", "t_end");
    const int nsnap = cfg.i("time", "n_snaps");
    const float M0 = cfg.f("thermo", "M0"), sig0 = cfg.f("thermo", "sigma0");
    const float eps4 = cfg.f("thermo", "eps4", 0.04f);
    const float D = cfg.f("thermo", "D"), m_liq = cfg.f("thermo", "m_liq");
    const float k_part = cfg.f("thermo", "k_part"), T_melt = cfg.f("thermo", "T_melt");
    const float L = cfg.f("thermo", "L");
    const float c0 = cfg.f("initial", "c0"), dT = cfg.f("initial", "undercooling");
    const int seed_r = cfg.i("initial", "seed_radius", 4);
    const bool elastic_enabled = cfg.b("thermo", "elastic_enabled", false);
    const float elastic_modulus = cfg.f("thermo", "elastic_modulus", 0.0f);
    const float eigenstrain_coeff = cfg.f("thermo", "eigenstrain_coeff", 0.0f);
    const int elastic_iters = cfg.i("thermo", "elastic_smooth_iters", 8);
    const float c_ref = cfg.f("thermo", "c_ref", c0);

    std::printf("=== GPK Solidification (MI355X) ===\n");
    std::printf("Grid %d×%d×%d, dx=%.2f µm\n", Nz, Ny, Nx, dx * 1e6f);
    auto thermo = std::make_shared<SolidificationThermo>(M0, sig0, eps4, m_liq, k_part, T_melt, L);
    thermo->stochastic = cfg.b("thermo", "stochastic", false);
    thermo->noise_scale = cfg.f("thermo", "noise_scale", 1.0f);
    thermo->elastic_enabled = elastic_enabled;
    thermo->elastic_modulus = elastic_modulus;
    thermo->eigenstrain_coeff = eigenstrain_coeff;
    thermo->c_ref = c_ref;
    thermo->elastic_smooth_iters = elastic_iters;
    thermo->solute_rejection = cfg.b("thermo", "solute_rejection", false);
    thermo->k_partition = k_part;
    if (thermo->solute_rejection) std::printf("Solute partitioning enabled, k = %.3f\n", double(k_part));
    GPKGrid grid(Nz, Ny, Nx, dx, thermo, D, cfg.f("geometry", "sigma_cells", 2.f), precision_of(cfg),
                 cfg.i("gpu", "device", 0));
    const float T_liq = T_melt + m_liq * c0;
    std::fill(grid.c.begin(), grid.c.end(), c0);
    std::fill(grid.T.begin(), grid.T.end(), T_liq - dT);
    const float R_star = sig0 * T_melt / (L * dT);
    std::printf("Gibbs-Thomson R* = %.3f µm\n", R_star * 1e6f);
    std::printf("Seed radius      = %.1f µm (must be > R*)\n\n", seed_r * dx * 1e6f);
    if (thermo->stochastic)
        std::printf("Stochastic kinetics enabled, noise_scale = %.3f\n", double(thermo->noise_scale));
    if (elastic_enabled)
        std::printf("Elastic coupling enabled, modulus = %.3e J/m^3, coeff = %.3e\n\n", double(elastic_modulus),
                    double(eigenstrain_coeff));
    const int cz = Nz / 2, cy = Ny / 2, cx = Nx / 2;
    for (int dz = -seed_r; dz <= seed_r; ++dz)
        for (int dy = -seed_r; dy <= seed_r; ++dy)
            for (int dxs = -seed_r; dxs <= seed_r; ++dxs) {
                if (dz * dz + dy * dy + dxs * dxs > seed_r * seed_r) continue;
                const int zz = cz + dz, yy = cy + dy, xx = cx + dxs;
                if (zz < 0 || zz >= Nz || yy < 0 || yy >= Ny || xx < 0 || xx >= Nx) continue;
                const int i = grid.idx(zz, yy, xx);
                grid.phi[i] = 1;
                grid.c[i] = c0 * k_part;
            }
    grid.mark_dirty();
    Writer writer(cfg);
    const bool write_initial = cfg.b("output", "write_initial", false);
    const int off = write_initial ? 1 : 0;
    if (write_initial) writer.write(grid, 0, grid.time);
    const double snap_dt = t_end / nsnap;
    std::printf("%-10s %-14s %-14s %-10s\n", "Time[ms]", "Solid frac", "Solute max", "Steps");
    const auto t0 = std::chrono::high_resolution_clock::now();
    int total = 0;
    for (int snap = 0; snap < nsnap; ++snap) {
        const int steps = grid.run(snap_dt * (snap + 1), -1.f);
        total += steps;
        grid.sync_host();
        const float sf = grid.solid_fraction();
        const float cmax = *std::max_element(grid.c.begin(), grid.c.end());
        std::printf("%-10.3f %-14.4f %-14.4f %-10d\n", grid.time * 1e3, sf, cmax, steps);
        writer.write(grid, snap + off, grid.time);
    }
    const double wall = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    std::printf("\nWall time: %.2f s  (%d steps, %.2f ms/step)\n", wall, total,
                total ? 1e3 * wall / total : 0.0);
    return 0;
}

// demos/gpk.cpp:397-426: n_nuclei discs of radius r_nuc, centres and
// orientations from the driver's generator, written into matrix cells only
void rex_nucleate(GPKGrid& grid, int n_nuclei, int r_nuc, std::mt19937& rng, int& next_gid) {
    const int Nz = grid.Nz, Ny = grid.Ny, Nx = grid.Nx;
    const bool is2d = Nz == 1;
    std::uniform_int_distribution<int> rz(is2d ? 0 : r_nuc, is2d ? 0 : (Nz - r_nuc - 1));
    std::uniform_int_distribution<int> ry(r_nuc, Ny - r_nuc - 1);
    std::uniform_int_distribution<int> rx(r_nuc, Nx - r_nuc - 1);
    std::uniform_real_distribution<float> rg(0.f, float(M_PI) * 0.5f);
    for (int n = 0; n < n_nuclei; ++n) {
        const int cz = rz(rng);
        const int cy = ry(rng);
        const int cx = rx(rng);
        const float gval = rg(rng);
        const int gid = ++next_gid;
        for (int dz = -r_nuc; dz <= r_nuc; ++dz)
            for (int dy = -r_nuc; dy <= r_nuc; ++dy)
                for (int ddx = -r_nuc; ddx <= r_nuc; ++ddx) {
                    if (dz * dz + dy * dy + ddx * ddx > r_nuc * r_nuc) continue;
                    const int z = cz + dz, y = cy + dy, x = cx + ddx;
                    if (z < 0 || z >= Nz || y < 0 || y >= Ny || x < 0 || x >= Nx) continue;
                    if (is2d && dz != 0) continue;
                    const int i = grid.idx(z, y, x);
                    if (grid.phi[i] == 0) {
                        grid.phi[i] = gid;
                        grid.g[i] = gval;
                        grid.alpha[i] = 1.f;
                    }
                }
    }
}

// demos/gpk.cpp:428-475: per step, every matrix cell away from the border
// seeds a nucleus with probability 1 - exp(-Jdot V dt); needs the labels on
// the host, so the mirrors are synchronised first and re-uploaded if seeded
int rex_stochastic_nucleation(GPKGrid& grid, float Jdot, float dt, int r_nuc, std::mt19937& rng,
                              int& next_gid) {
    if (Jdot <= 0.f || dt <= 0.f) return 0;
    const int Nz = grid.Nz, Ny = grid.Ny, Nx = grid.Nx;
    const float Vcell = grid.dx * grid.dx * grid.dx;
    const float P = 1.f - std::exp(-Jdot * Vcell * dt);
    if (P <= 0.f) return 0;
    grid.sync_host();
    std::uniform_real_distribution<float> u01(0.f, 1.f);
    std::uniform_real_distribution<float> rg(0.f, float(M_PI) * 0.5f);
    int added = 0;
    auto try_seed = [&](int z, int y, int x) {
        if (grid.phi[grid.idx(z, y, x)] != 0) return;
        if (u01(rng) > P) return;
        const int gid = ++next_gid;
        const float gval = rg(rng);
        for (int dz = -r_nuc; dz <= r_nuc; ++dz)
            for (int dy = -r_nuc; dy <= r_nuc; ++dy)
                for (int ddx = -r_nuc; ddx <= r_nuc; ++ddx) {
                    if (dz * dz + dy * dy + ddx * ddx > r_nuc * r_nuc) continue;
                    const int zz = z + dz, yy = y + dy, xx = x + ddx;
                    if (zz < 0 || zz >= Nz || yy < 0 || yy >= Ny || xx < 0 || xx >= Nx) continue;
                    const int j = grid.idx(zz, yy, xx);
                    if (grid.phi[j] == 0) {
                        grid.phi[j] = gid;
                        grid.g[j] = gval;
                        grid.alpha[j] = 1.f;
                    }
                }
        ++added;
    };
    if (Nz == 1) {
        for (int y = r_nuc; y < Ny - r_nuc; ++y)
            for (int x = r_nuc; x < Nx - r_nuc; ++x) try_seed(0, y, x);
    } else {
        for (int z = r_nuc; z < Nz - r_nuc; ++z)
            for (int y = r_nuc; y < Ny - r_nuc; ++y)
                for (int x = r_nuc; x < Nx - r_nuc; ++x) try_seed(z, y, x);
    }
    if (added) grid.mark_dirty();
    return added;
}

// demos/gpk.cpp:477-555
int run_recrystallisation(const Deck& cfg) {
    const int Nz = cfg.i("grid", "Nz"), Ny = cfg.i("grid", "Ny"), Nx = cfg.i("grid", "Nx");
    const float dx = cfg.f("grid", "dx");
    const double t_end = cfg.d("time", "t_end");
    const int n_snaps = cfg.i("time", "n_snaps");
    const float M0 = cfg.f("thermo", "M0"), sig0 = cfg.f("thermo", "sigma0");
    const float E_stored = cfg.f("thermo", "E_stored");
    const int N_NUC = cfg.i("nucleation", "n_nuclei");
    const int R_NUC = cfg.i("nucleation", "nucleus_r");
    const float Jdot = cfg.f("nucleation", "Jdot", 0.f);
    const int seed = cfg.i("nucleation", "seed", 42);

    std::printf("=== GPK Recrystallisation (MI355X) ===\n");
    std::printf("Grid %d×%d×%d, dx=%.1f µm\n", Nz, Ny, Nx, dx * 1e6f);
    std::printf("E_stored = %.2e J/m³\n\n", double(E_stored));
    const float R_star = sig0 / E_stored;
    std::printf("Critical R* = %.3f µm (nucleus radius = %d cells = %.1f µm)\n\n", R_star * 1e6f, R_NUC,
                R_NUC * dx * 1e6f);
    auto thermo = std::make_shared<RecrystallisationThermo>(M0, sig0, E_stored);
    GPKGrid grid(Nz, Ny, Nx, dx, thermo, 0.f, cfg.f("geometry", "sigma_cells", 2.f), precision_of(cfg),
                 cfg.i("gpu", "device", 0));
    std::fill(grid.phi.begin(), grid.phi.end(), 0);
    std::fill(grid.alpha.begin(), grid.alpha.end(), 1.f);
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    std::uniform_real_distribution<float> rg_dist(0.f, float(M_PI) * 0.5f);
    for (auto& v : grid.g) v = rg_dist(rng);
    int next_gid = 0;
    rex_nucleate(grid, N_NUC, R_NUC, rng, next_gid);
    grid.mark_dirty();
    std::printf("Nucleated %d rex grains (radius %d cells), Jdot = %.3e 1/(m^3 s)\n\n", N_NUC, R_NUC,
                double(Jdot));
    Writer writer(cfg);
    const bool write_initial = cfg.b("output", "write_initial", false);
    const int off = write_initial ? 1 : 0;
    if (write_initial) writer.write(grid, 0, grid.time);
    const double snap_dt = t_end / n_snaps;
    std::printf("%-10s %-12s %-14s %-10s\n", "Time[s]", "Rex frac", "ln(-ln(1-X))", "N_grains");
    const auto t0 = std::chrono::high_resolution_clock::now();
    for (int snap = 0; snap < n_snaps; ++snap) {
        const double t_target = snap_dt * (snap + 1);
        while (grid.time < t_target) {
            const float dt = grid.step(-1.f);
            rex_stochastic_nucleation(grid, Jdot, dt, R_NUC, rng, next_gid);
        }
        grid.sync_host();
        int rex_cnt = 0;
        for (auto v : grid.phi) if (v != 0) ++rex_cnt;
        const float X = float(rex_cnt) / float(grid.N());
        const float avrami = (X > 0.f && X < 1.f) ? std::log(-std::log(1.f - X)) : (X >= 1.f ? 99.f : -99.f);
        std::printf("%-10.4f %-12.4f %-14.3f %-10d\n", grid.time, X, avrami, grid.n_grains());
        writer.write(grid, snap + off, grid.time);
        if (X > 0.99f) {
            std::printf("Fully recrystallised.\n");
            break;
        }
    }
    const double wall = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    std::printf("\nWall time: %.2f s\n", wall);
    return 0;
}

// demos/gpk.cpp:693-701
std::vector<float> parse_csv_floats(const std::string& s) {
    std::vector<float> out;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ','))
        if (!tok.empty()) out.push_back(std::stof(tok));
    return out;
}

// demos/gpk.cpp:703-797
int run_precipitation(const Deck& cfg) {
    const int Nz = cfg.i("grid", "Nz"), Ny = cfg.i("grid", "Ny"), Nx = cfg.i("grid", "Nx");
    const float dx = cfg.f("grid", "dx");
    const double t_end = cfg.d("time", "t_end");
    const int nsnap = cfg.i("time", "n_snaps");
    const float M0 = cfg.f("thermo", "M0"), sig0 = cfg.f("thermo", "sigma0");
    const float D = cfg.f("thermo", "D", 0.f);
    const int n_phases = cfg.i("thermo", "n_phases");
    const float T_ref = cfg.f("thermo", "T_ref", 900.f);
    const std::vector<float> a = parse_csv_floats(cfg.str("thermo", "a"));
    const std::vector<float> b = parse_csv_floats(cfg.str("thermo", "b"));
    const std::vector<float> c2 = parse_csv_floats(cfg.str("thermo", "c2"));
    const std::vector<float> dT = parse_csv_floats(cfg.str("thermo", "dT"));
    const float c0 = cfg.f("initial", "c0"), T0 = cfg.f("initial", "T0");
    const int n_seeds = cfg.i("initial", "n_seeds");
    const int seed_r = cfg.i("initial", "seed_radius", 2);
    const int seed = cfg.i("initial", "seed", 123);

    std::printf("=== GPK Multi-phase Precipitation (MI355X) ===\n");
    std::printf("Grid %d×%d×%d, phases=%d, dx=%.2f µm\n\n", Nz, Ny, Nx, n_phases, dx * 1e6f);
    auto thermo = std::make_shared<PrecipitationThermo>(M0, sig0, n_phases, a, b, c2, dT, T_ref);
    GPKGrid grid(Nz, Ny, Nx, dx, thermo, D, cfg.f("geometry", "sigma_cells", 2.f), precision_of(cfg),
                 cfg.i("gpu", "device", 0));
    std::fill(grid.phi.begin(), grid.phi.end(), 0);
    std::fill(grid.alpha.begin(), grid.alpha.end(), 1.f);
    std::fill(grid.c.begin(), grid.c.end(), c0);
    std::fill(grid.T.begin(), grid.T.end(), T0);
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    std::uniform_int_distribution<int> rz(seed_r, std::max(seed_r, Nz - seed_r - 1));
    std::uniform_int_distribution<int> ry(seed_r, std::max(seed_r, Ny - seed_r - 1));
    std::uniform_int_distribution<int> rx(seed_r, std::max(seed_r, Nx - seed_r - 1));
    std::uniform_int_distribution<int> rp(1, std::max(1, n_phases - 1));
    for (int n = 0; n < n_seeds; ++n) {
        const int cz = Nz > 1 ? rz(rng) : 0;
        const int cy = ry(rng);
        const int cx = rx(rng);
        const int p = rp(rng);
        bool overlap = false;
        const int z_guard = Nz > 1 ? seed_r + 1 : 0;
        for (int dz = -z_guard; dz <= z_guard && !overlap; ++dz)
            for (int dy = -seed_r - 1; dy <= seed_r + 1 && !overlap; ++dy)
                for (int ddx = -seed_r - 1; ddx <= seed_r + 1 && !overlap; ++ddx) {
                    const int zz = cz + dz, yy = cy + dy, xx = cx + ddx;
                    if (zz < 0 || zz >= Nz || yy < 0 || yy >= Ny || xx < 0 || xx >= Nx) continue;
                    if (grid.phi[grid.idx(zz, yy, xx)] != 0) overlap = true;
                }
        if (overlap) continue;
        const int z_radius = Nz > 1 ? seed_r : 0;
        for (int dz = -z_radius; dz <= z_radius; ++dz)
            for (int dy = -seed_r; dy <= seed_r; ++dy)
                for (int ddx = -seed_r; ddx <= seed_r; ++ddx) {
                    if (dz * dz + dy * dy + ddx * ddx > seed_r * seed_r) continue;
                    const int zz = cz + dz, yy = cy + dy, xx = cx + ddx;
                    if (zz < 0 || zz >= Nz || yy < 0 || yy >= Ny || xx < 0 || xx >= Nx) continue;
                    grid.phi[grid.idx(zz, yy, xx)] = p;
                }
    }
    grid.mark_dirty();
    Writer writer(cfg);
    const bool write_initial = cfg.b("output", "write_initial", false);
    const int off = write_initial ? 1 : 0;
    if (write_initial) writer.write(grid, 0, grid.time);
    const double snap_dt = t_end / nsnap;
    std::printf("%-10s %-12s %-12s %-12s\n", "Time[s]", "Phase0", "Phase1", "Phase2+");
    const auto t0 = std::chrono::high_resolution_clock::now();
    for (int snap = 0; snap < nsnap; ++snap) {
        grid.run(snap_dt * (snap + 1), -1.f);
        grid.sync_host();
        std::unordered_map<int32_t, int> cnt;
        for (auto p : grid.phi) cnt[p]++;
        const float f0 = float(cnt[0]) / float(grid.N());
        const float f1 = float(cnt[1]) / float(grid.N());
        const float f2p = std::max(0.f, 1.f - f0 - f1);
        std::printf("%-10.4f %-12.4f %-12.4f %-12.4f\n", grid.time, f0, f1, f2p);
        writer.write(grid, snap + off, grid.time);
    }
    const double wall = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    std::printf("\nWall time: %.2f s\n", wall);
    return 0;
}

// ---- precipitate break-up sweep (demos/gpk.cpp:1065-1275) ----
// Per spacing d of [breakup] d_list: a periodic slab of thickness h (z),
// split into grains 1 = [0, d) and 2 = [d, 2d) along x (Nx = 2d, two grain
// boundaries per period), beta (label 0) above and below; pure capillary
// flow (AnisoOstwaldThermo at c = c_sat, D = 0).  The run breaks when the
// alpha cells (label != 0) fall apart into at least two pieces of
// min_piece cells or more.  The labels are read between steps, so the host
// mirrors are refreshed from the device first.

// periodic face-connected pieces of the alpha phase with at least min_size
// cells (demos/gpk.cpp:975-1002)
int alpha_pieces(const GPKGrid& g, int min_size) {
    const int Nz = g.Nz, Ny = g.Ny, Nx = g.Nx, N = g.N();
    std::vector<uint8_t> seen(size_t(N), 0);
    std::vector<int> queue;
    queue.reserve(size_t(N));
    int big = 0;
    for (int s = 0; s < N; ++s) {
        if (g.phi[size_t(s)] == 0 || seen[size_t(s)]) continue;
        queue.clear();
        queue.push_back(s);
        seen[size_t(s)] = 1;
        for (size_t head = 0; head < queue.size(); ++head) {
            int z, y, x;
            g.decode(queue[head], z, y, x);
            const int nb[6] = {g.idx(z, y, x == 0 ? Nx - 1 : x - 1), g.idx(z, y, x == Nx - 1 ? 0 : x + 1),
                               g.idx(z, y == 0 ? Ny - 1 : y - 1, x), g.idx(z, y == Ny - 1 ? 0 : y + 1, x),
                               g.idx(z == 0 ? Nz - 1 : z - 1, y, x), g.idx(z == Nz - 1 ? 0 : z + 1, y, x)};
            for (int j : nb)
                if (g.phi[size_t(j)] != 0 && !seen[size_t(j)]) {
                    seen[size_t(j)] = 1;
                    queue.push_back(j);
                }
        }
        if (int(queue.size()) >= min_size) ++big;
    }
    return big;
}

// label slice as comma-separated rows: x-z at y = Ny/2 (rows z) or x-y at
// z = Nz/2 (rows y) (demos/gpk.cpp:1006-1030)
void label_slice_csv(const GPKGrid& g, const std::string& path, bool xz) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return;
    const int rows = xz ? g.Nz : g.Ny;
    for (int r = 0; r < rows; ++r)
        for (int x = 0; x < g.Nx; ++x) {
            const int i = xz ? g.idx(r, g.Ny / 2, x) : g.idx(g.Nz / 2, r, x);
            std::fprintf(f, "%d%c", g.phi[size_t(i)], x == g.Nx - 1 ? '\n' : ',');
        }
    std::fclose(f);
}

// whole grid, one row per cell, in the microstructure CSV layout
// (demos/gpk.cpp:1036-1062)
void grid_csv(const GPKGrid& g, const std::string& path, double t, int snap) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return;
    std::vector<char> buf(size_t(8) << 20);
    std::setvbuf(f, buf.data(), _IOFBF, buf.size());
    std::fprintf(f, "# GPK Microstructure CSV\n# snap=%d  time=%g\n# Nz=%d  Ny=%d  Nx=%d\n# dx=%g [m]\n", snap, t,
                 g.Nz, g.Ny, g.Nx, double(g.dx));
    std::fprintf(f, "# cols: x[m] y[m] z[m] grain_id alpha composition temperature orientation\n");
    const double h = double(g.dx);
    for (int z = 0; z < g.Nz; ++z)
        for (int y = 0; y < g.Ny; ++y)
            for (int x = 0; x < g.Nx; ++x) {
                const size_t i = size_t(g.idx(z, y, x));
                std::fprintf(f, "%.6e %.6e %.6e %d %.6f %.6f %.2f %.6f\n", x * h, y * h, z * h, g.phi[i],
                             g.alpha[i], g.c[i], g.T[i], g.g[i]);
            }
    std::fclose(f);
}

int run_precipitate_breakup(const Deck& cfg) {
    const int Ny = cfg.i("grid", "Ny", 96), Nz = cfg.i("grid", "Nz", 48);
    const float dx = cfg.f("grid", "dx", 5.0e-9f);
    const int nsnap = cfg.i("time", "n_snaps", 30);
    const float M0 = cfg.f("thermo", "M0", 1.0e-9f);
    const float sig_broad = cfg.f("thermo", "sigma_broad", 0.091f);
    const float sig_side = cfg.f("thermo", "sigma_side", 0.811f);
    const float c_sat = cfg.f("thermo", "c_sat", 0.15f);
    const float Omega = cfg.f("thermo", "Omega", 5.0e8f);
    const int h = cfg.i("breakup", "h_cells", 16);
    const int n_steps = cfg.i("breakup", "n_steps", 2000);
    const float min_frac = cfg.f("breakup", "min_piece_frac", 0.25f);
    const std::vector<float> d_list = parse_csv_floats(cfg.str("breakup", "d_list"));
    if (d_list.empty()) {
        std::fprintf(stderr, "breakup.d_list is empty\n");
        return 1;
    }
    if (Nz < 2) {
        std::fprintf(stderr, "precipitate_breakup is 3-D: set grid.Nz > 1\n");
        return 1;
    }
    const std::string dir = cfg.str("output", "dir", "output/precipitate_breakup");
    // sigma(z) = sigma0 sqrt(bzz) = sigma_broad, sigma(x) = sigma(y) = sigma0 = sigma_side
    const float bzz = (sig_broad / sig_side) * (sig_broad / sig_side);
    std::printf("=== GPK Precipitate Breakup 3-D (MI355X) ===\n");
    std::printf("Nz=%d Ny=%d (Nx=2d per d)  |  h=%d cells  |  %d steps per run\n", Nz, Ny, h, n_steps);
    std::printf("sigma_broad=%.3f  sigma_side=%.3f J/m^2  -> sigma0=%.3f bxx=byy=1 bzz=%.5f\n\n", double(sig_broad),
                double(sig_side), double(sig_side), double(bzz));
    auto thermo = std::make_shared<AnisoOstwaldThermo>(M0, sig_side, 1.f, 1.f, bzz, c_sat, Omega);
    mkdir_p(dir);
    FILE* sf = std::fopen((dir + "/summary.csv").c_str(), "w");
    if (!sf) throw std::runtime_error("cannot write " + dir + "/summary.csv");
    std::fprintf(sf, "d_cells,h_cells,dh,max_pieces,result\n");
    std::printf("%-8s %-8s %-12s %-10s\n", "d[cell]", "d/h", "max_pieces", "result");
    double dh_break = 1e30, dh_stable = 0.0;
    const int snap_every = std::max(1, n_steps / nsnap);
    const int ts_every = std::max(1, n_steps / 10);
    auto name = [&](int d, const char* what) {
        char b[64];
        std::snprintf(b, sizeof(b), "/d%03d_%s.csv", d, what);
        return dir + b;
    };
    for (const float df : d_list) {
        const int d = int(std::lround(df));
        const float dh = float(d) / float(h);
        GPKGrid grid(Nz, Ny, 2 * d, dx, thermo, 0.f, cfg.f("geometry", "sigma_cells", 2.f), precision_of(cfg),
                     cfg.i("gpu", "device", 0));
        std::fill(grid.phi.begin(), grid.phi.end(), 0);
        std::fill(grid.alpha.begin(), grid.alpha.end(), 1.f);
        std::fill(grid.g.begin(), grid.g.end(), 0.f);
        std::fill(grid.c.begin(), grid.c.end(), c_sat);
        std::fill(grid.T.begin(), grid.T.end(), 973.f);
        const int zlo = Nz / 2 - h / 2;
        for (int z = zlo; z < zlo + h; ++z)
            for (int y = 0; y < Ny; ++y)
                for (int x = 0; x < 2 * d; ++x) grid.phi[size_t(grid.idx(z, y, x))] = x < d ? 1 : 2;
        grid.mark_dirty();
        const long alpha0 = long(std::count_if(grid.phi.begin(), grid.phi.end(), [](int32_t v) { return v != 0; }));
        const int min_piece = std::max(8, int(std::lround(min_frac * double(alpha0) / 2)));
        label_slice_csv(grid, name(d, "xz_t0"), true);
        label_slice_csv(grid, name(d, "xy_t0"), false);
        grid_csv(grid, name(d, "3d_t0"), 0.0, 0);
        grid_csv(grid, name(d, "3d_ts0000"), 0.0, 0);
        // the most pieces seen (and that state), and the last state that
        // still held 80% of the alpha cells (checked every 4th step)
        int max_pieces = 1;
        std::vector<int32_t> best_phi, grooved_phi;
        double best_time = 0.0, grooved_time = 0.0;
        const long alpha_floor = long(0.80 * double(alpha0));
        for (int s = 0; s < n_steps; ++s) {
            grid.step(-1.f);
            const bool vol = s % 4 == 0;
            const bool pieces = (s + 1) % snap_every == 0 || s == n_steps - 1;
            const bool dump = (s + 1) % ts_every == 0 || s == n_steps - 1;
            if (dump) grid.sync_host();
            else if (vol || pieces) grid.sync_labels();
            if (vol) {
                const long cur = long(std::count_if(grid.phi.begin(), grid.phi.end(), [](int32_t v) { return v != 0; }));
                if (cur >= alpha_floor) {
                    grooved_phi = grid.phi;
                    grooved_time = grid.time;
                }
            }
            if (pieces) {
                const int big = alpha_pieces(grid, min_piece);
                if (big > max_pieces) {
                    max_pieces = big;
                    best_phi = grid.phi;
                    best_time = grid.time;
                }
            }
            if (dump) {
                char ts[32];
                std::snprintf(ts, sizeof(ts), "3d_ts%04d", s + 1);
                grid_csv(grid, name(d, ts), grid.time, s + 1);
            }
        }
        grid.sync_host();
        label_slice_csv(grid, name(d, "xz_tf"), true);
        label_slice_csv(grid, name(d, "xy_tf"), false);
        // the severed state if there was one, else the last grooved state,
        // else the final one (labels only; the other fields stay final)
        const std::vector<int32_t> final_phi = grid.phi;
        const bool have_best = !best_phi.empty(), have_groove = !grooved_phi.empty();
        const double disp_time = have_best ? best_time : (have_groove ? grooved_time : grid.time);
        if (have_best) grid.phi = best_phi;
        else if (have_groove) grid.phi = grooved_phi;
        label_slice_csv(grid, name(d, "xz_tb"), true);
        label_slice_csv(grid, name(d, "xy_tb"), false);
        grid_csv(grid, name(d, "3d_tb"), disp_time, 1);
        grid.phi = final_phi;
        const bool broke = max_pieces >= 2;
        const char* res = broke ? "BREAK" : "stable";
        std::printf("%-8d %-8.2f %-12d %-10s\n", d, double(dh), max_pieces, res);
        std::fprintf(sf, "%d,%d,%.3f,%d,%s\n", d, h, double(dh), max_pieces, res);
        if (broke) dh_break = std::min(dh_break, double(dh));
        else dh_stable = std::max(dh_stable, double(dh));
    }
    std::fclose(sf);
    std::printf("\n");
    if (dh_break < 1e29 && dh_stable > 0.0) {
        std::printf("Largest stable d/h    = %.2f\n", dh_stable);
        std::printf("Smallest breaking d/h = %.2f\n", dh_break);
        std::printf("==> Critical aspect ratio (d/h)* ~ %.2f\n", 0.5 * (dh_break + dh_stable));
    } else if (dh_break < 1e29) {
        std::printf("All tested ratios broke up; (d/h)* < %.2f\n", dh_break);
    } else {
        std::printf("No breakup observed; (d/h)* > %.2f\n", dh_stable);
    }
    std::printf("Snapshots + summary written to %s\n", dir.c_str());
    return 0;
}

} // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "Usage: %s <input.ini>\n", argv[0]);
        return 1;
    }
    try {
        Deck cfg(argv[1]);
        const std::string model = cfg.str("run", "model");
        if (model == "grain_growth") return run_grain_growth(cfg);
        if (model == "solidification") return run_solidification(cfg);
        if (model == "recrystallisation") return run_recrystallisation(cfg);
        if (model == "precipitation") return run_precipitation(cfg);
        if (model == "precipitate_breakup") return run_precipitate_breakup(cfg);
        std::fprintf(stderr, "Unknown or unsupported model '%s' (this build runs grain_growth, "
                             "solidification, recrystallisation, precipitation and precipitate_breakup)\n",
                     model.c_str());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}
