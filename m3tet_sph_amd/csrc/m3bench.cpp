// m3bench: standalone C++ harness of the C ABI (include/m3stokes.h).  It
// reads a solve case written by m3tet_sph_amd.api.export_case (the
// m3_mesh_desc arrays, UBC.ifix / UBC.vfix, nodal Visc / Dens as .npy files,
// sizes and PHYSICS / NUMSCALE scalars in case.txt), runs the whole Stokes
// solve m3_flow3d_patera_sph (the flow3d_patera_sph.m:1-9 call of the
// reference's time loop) W + K times and prints one JSON line with the rate,
// the iteration counts and the residuals of the last solve.  With --out it
// writes the solution U (3*nnod, rotated frame) and P (nVnod) as .npy files.
//
//   m3bench CASE_DIR [--steps K] [--warmup W] [--precision fp64|mixed]
//                    [--rtol R] [--out DIR] [--device D]
//
// No Python, no MATLAB: the ABI is all it uses.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/m3stokes.h"

namespace {

struct Npy {
  std::string descr;
  std::vector<int64_t> shape;
  std::vector<char> data;
  int64_t count() const {
    int64_t n = 1;
    for (int64_t s : shape) n *= s;
    return n;
  }
};

// .npy reader for little-endian C-ordered arrays (format versions 1-3)
Npy read_npy(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  char magic[6];
  f.read(magic, 6);
  if (std::memcmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error(path + ": not a .npy file");
  unsigned char ver[2];
  f.read(reinterpret_cast<char *>(ver), 2);
  uint32_t hlen = 0;
  if (ver[0] == 1) {
    unsigned char b[2];
    f.read(reinterpret_cast<char *>(b), 2);
    hlen = b[0] | (b[1] << 8);
  } else {
    unsigned char b[4];
    f.read(reinterpret_cast<char *>(b), 4);
    hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
  }
  std::string h(hlen, ' ');
  f.read(&h[0], hlen);
  Npy a;
  const size_t d = h.find("'descr'");
  const size_t q1 = h.find('\'', h.find(':', d) + 1), q2 = h.find('\'', q1 + 1);
  a.descr = h.substr(q1 + 1, q2 - q1 - 1);
  if (h.find("'fortran_order': False") == std::string::npos)
    throw std::runtime_error(path + ": Fortran-ordered arrays are not read");
  const size_t s0 = h.find('(', h.find("'shape'")), s1 = h.find(')', s0);
  std::stringstream ss(h.substr(s0 + 1, s1 - s0 - 1));
  std::string tok;
  while (std::getline(ss, tok, ','))
    if (tok.find_first_not_of(" ") != std::string::npos) a.shape.push_back(std::stoll(tok));
  const int isz = std::stoi(a.descr.substr(2));
  a.data.resize((size_t)a.count() * isz);
  f.read(a.data.data(), (std::streamsize)a.data.size());
  if (!f) throw std::runtime_error(path + ": truncated");
  return a;
}

template <class T>
std::vector<T> load(const std::string &dir, const std::string &name, const char *descr) {
  Npy a = read_npy(dir + "/" + name + ".npy");
  if (a.descr != descr && !(a.descr.size() == 3 && a.descr[0] == '|' && std::string(descr).substr(1) == a.descr.substr(1)))
    throw std::runtime_error(name + ": dtype " + a.descr + ", expected " + descr);
  std::vector<T> v((size_t)a.count());
  std::memcpy(v.data(), a.data.data(), a.data.size());
  return v;
}

void write_npy(const std::string &path, const std::vector<double> &v) {
  std::string h = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + std::to_string(v.size()) + ",), }";
  while ((10 + h.size() + 1) % 64) h += ' ';
  h += '\n';
  std::ofstream f(path, std::ios::binary);
  f.write("\x93NUMPY\x01\x00", 8);
  const uint16_t hl = (uint16_t)h.size();
  f.write(reinterpret_cast<const char *>(&hl), 2);
  f.write(h.data(), (std::streamsize)h.size());
  f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}

std::map<std::string, std::string> read_case(const std::string &dir) {
  std::ifstream f(dir + "/case.txt");
  if (!f) throw std::runtime_error("cannot open " + dir + "/case.txt");
  std::map<std::string, std::string> kv;
  std::string k, v;
  while (f >> k >> v) kv[k] = v;
  return kv;
}

void check(m3_ctx *c, int rc, const char *what) {
  if (rc) {
    std::fprintf(stderr, "m3bench: %s failed (%d): %s\n", what, rc, c ? m3_last_error(c) : "");
    std::exit(1);
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2 || std::string(argv[1]) == "--help") {
    std::fprintf(stderr,
                 "usage: m3bench CASE_DIR [--steps K] [--warmup W] [--precision fp64|mixed] [--rtol R] "
                 "[--out DIR] [--device D]\n");
    return argc < 2 ? 2 : 0;
  }
  const std::string dir = argv[1];
  int steps = 5, warmup = 1, device = 0, prec = M3_PREC_MIXED;
  double rtol = 1e-6;
  std::string out;
  for (int i = 2; i + 1 < argc; i += 2) {
    const std::string a = argv[i], v = argv[i + 1];
    if (a == "--steps") steps = std::atoi(v.c_str());
    else if (a == "--warmup") warmup = std::atoi(v.c_str());
    else if (a == "--device") device = std::atoi(v.c_str());
    else if (a == "--rtol") rtol = std::atof(v.c_str());
    else if (a == "--out") out = v;
    else if (a == "--precision") prec = v == "fp64" ? M3_PREC_FP64 : M3_PREC_MIXED;
    else {
      std::fprintf(stderr, "m3bench: unknown option %s\n", a.c_str());
      return 2;
    }
  }
  try {
    auto kv = read_case(dir);
    const int nmg = std::stoi(kv.at("nmg"));
    m3_mesh_desc d{};
    d.nmg = nmg;
    std::vector<std::vector<uint32_t>> el(nmg);
    std::vector<std::vector<int64_t>> ip(nmg);
    std::vector<std::vector<int32_t>> ic(nmg);
    std::vector<std::vector<double>> iw(nmg);
    for (int l = 0; l < nmg; ++l) {
      const std::string s = std::to_string(l);
      d.nel[l] = std::stoll(kv.at("nel_" + s));
      d.nnod[l] = std::stoll(kv.at("nnod_" + s));
      d.nVnod[l] = std::stoll(kv.at("nVnod_" + s));
      el[l] = load<uint32_t>(dir, "EL2NOD_" + s, "<u4");
      d.EL2NOD[l] = el[l].data();
      if (l < nmg - 1) {
        ip[l] = load<int64_t>(dir, "Ic2f_ptr_" + s, "<i8");
        ic[l] = load<int32_t>(dir, "Ic2f_col_" + s, "<i4");
        iw[l] = load<double>(dir, "Ic2f_val_" + s, "<f8");
        d.Ic2f_ptr[l] = ip[l].data();
        d.Ic2f_col[l] = ic[l].data();
        d.Ic2f_val[l] = iw[l].data();
      }
    }
    auto X = load<double>(dir, "GCOORD", "<f8"), S = load<double>(dir, "GCOORD_SPH", "<f8");
    d.GCOORD = X.data();
    d.GCOORD_SPH = S.data();
    std::vector<int32_t> cls[4];
    for (int k = 0; k < 4; ++k) {
      cls[k] = load<int32_t>(dir, "class_" + std::to_string(k), "<i4");
      d.els_class[k] = cls[k].data();
      d.nels_class[k] = (int64_t)cls[k].size();
    }
    auto E20 = load<uint32_t>(dir, "EL2NOD_cubic", "<u4");
    auto Xc = load<double>(dir, "GCOORD_cubic", "<f8");
    d.EL2NOD_cubic = E20.data();
    d.GCOORD_cubic = Xc.data();
    d.ncubic = std::stoll(kv.at("ncubic"));
    auto ifix = load<int64_t>(dir, "ifix", "<i8");
    auto vfix = load<double>(dir, "vfix", "<f8");
    auto Visc = load<double>(dir, "Visc", "<f8"), Dens = load<double>(dir, "Dens", "<f8");
    m3_phys ph{};
    ph.g = std::stod(kv.at("g"));
    ph.Bscale = std::stod(kv.at("Bscale"));
    ph.Dens0 = std::stod(kv.at("Dens0"));
    ph.has_minVisc = kv.count("minVisc");
    ph.has_maxVisc = kv.count("maxVisc");
    ph.minVisc = ph.has_minVisc ? std::stod(kv["minVisc"]) : 0.0;
    ph.maxVisc = ph.has_maxVisc ? std::stod(kv["maxVisc"]) : 0.0;
    ph.nip = 14;
    ph.type_Dmat = 0;

    m3_ctx *ctx = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    check(ctx, m3_create(&ctx, device), "m3_create");
    check(ctx, m3_set_mesh(ctx, &d), "m3_set_mesh");
    check(ctx, m3_set_bc(ctx, ifix.data(), vfix.data(), (int64_t)ifix.size()), "m3_set_bc");
    const double t_setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    m3_patera_opts o{};
    o.rtol_Pat = 0.8 * rtol;  // flow3d_patera_mg_sph_p.m:70-98 defaults otherwise
    o.rtol_U = 5e-3;
    o.itmax_U = 20;
    o.itmin_Pat = 3;
    o.itmax_Pat = 120;
    o.pc_Pat = 3;
    o.wght_MM = 1.0;
    o.precision = prec;
    o.warm_U = 0;
    o.rtol_U_final = 0.5 * rtol;
    o.itmax_U_final = 60;
    o.tol_Pat_abs = 0.0;
    const size_t n3 = 3 * (size_t)d.nnod[0], nV = (size_t)d.nVnod[0];
    std::vector<double> U(n3), P(nV);
    m3_info info{};
    auto solve = [&] {
      std::fill(U.begin(), U.end(), 0.0);
      std::fill(P.begin(), P.end(), 0.0);
      check(ctx, m3_flow3d_patera_sph(ctx, Visc.data(), Dens.data(), &ph, U.data(), P.data(), &o, &info),
            "m3_flow3d_patera_sph");
    };
    for (int w = 0; w < warmup; ++w) solve();
    auto t1 = std::chrono::steady_clock::now();
    for (int k = 0; k < steps; ++k) solve();
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count() / steps;
    double res[3];
    check(ctx, m3_residual(ctx, U.data(), P.data(), res), "m3_residual");
    std::printf(
        "{\"harness\": \"m3bench\", \"case\": \"%s\", \"precision\": \"%s\", \"velocity_dofs\": %zu, "
        "\"solves_per_s\": %.4f, \"ms_per_solve\": %.2f, \"steps\": %d, \"warmup\": %d, \"nitP\": %d, \"nitU\": %d, "
        "\"nitZ\": %d, \"rel_res_U\": %.3e, \"rel_res_P\": %.3e, \"setup_s\": %.2f}\n",
        dir.c_str(), prec == M3_PREC_FP64 ? "fp64" : "mixed", n3, 1.0 / sec, 1e3 * sec, steps, warmup, info.nitP,
        info.nitU, info.nitZ, res[0] / res[1], res[2] / info.norm_rP0, t_setup);
    if (!out.empty()) {
      write_npy(out + "/U.npy", U);
      write_npy(out + "/P.npy", P);
    }
    m3_destroy(ctx);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "m3bench: %s\n", e.what());
    return 1;
  }
  return 0;
}
