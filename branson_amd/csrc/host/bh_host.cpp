// bh_host.cpp -- the host side of the engine: Branson deck input, mesh,
// step physics and the replicated time loop (include/branson_host.h).
//
// Everything here is FP64 host arithmetic in the reference's operation order
// (built with -ffp-contract=off), because its results -- opacities, Fleck
// factors, per-cell source energies, photon counts, temperatures -- feed the
// device kernels and the conservation check.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../../include/branson_host.h"
#include "bh_xml.h"

namespace {

constexpr double kC = 299.792458;  // src/constants.h:17
constexpr double kA = 0.01372;     // src/constants.h:22

// ---------------------------------------------------------------------------
// Input (src/input.h:77-637)
// ---------------------------------------------------------------------------
struct Region {  // src/region.h:27-76
  uint32_t id = 0;
  double cv = 0, rho = 0, opA = 0, opB = 0, opC = 0, opS = 0, T_e = 0, T_r = 0;
  // Region::get_absorption_opacity (src/region.h:44-46)
  double absorption(double T) const { return opA + opB * std::pow(T, opC); }
};

struct Division {
  double start = 0, end = 0;
  uint32_t cells = 0;
};

struct Deck {
  double t_start = 0, t_stop = 0, dt = 0, t_mult = 0, dt_max = 0;
  uint64_t n_photons = 0;
  uint32_t seed = 0;
  std::vector<Division> div[3];
  std::map<uint32_t, uint32_t> region_map;  // z*1e6 + y*1e3 + x -> region ID
  std::vector<Region> regions;
  std::map<uint32_t, uint32_t> region_index;  // region ID -> index
  int32_t bc[6] = {0, 0, 0, 0, 0, 0};         // X_NEG .. Z_POS
  double T_source = 0;
};

// pugixml's text().as_double()/as_int() read a missing or empty node as 0.
double as_double(const bhx::Node* n, const char* tag) {
  const std::string t = n ? n->child_text(tag) : std::string();
  return t.empty() ? 0.0 : std::strtod(t.c_str(), nullptr);
}
long long as_ll(const bhx::Node* n, const char* tag) {
  const std::string t = n ? n->child_text(tag) : std::string();
  return t.empty() ? 0 : std::strtoll(t.c_str(), nullptr, 10);
}

bool parse_deck(const std::string& path, Deck& d, std::string& err) {
  auto root = bhx::parse_file(path, &err);
  if (!root) return false;
  if (root->name != "prototype") {
    err = "root element must be <prototype>";
    return false;
  }
  const bhx::Node* common = root->child("common");
  const bhx::Node* spatial = root->child("spatial");
  const bhx::Node* boundary = root->child("boundary");
  const bhx::Node* regions = root->child("regions");
  if (!common) return err = "'common' section not found!", false;
  if (!spatial) return err = "'spatial' section not found!", false;
  if (!boundary) return err = "'boundary' section not found!", false;
  if (!regions) return err = "'regions' section not found!", false;

  d.t_stop = as_double(common, "t_stop");
  d.dt = as_double(common, "dt_start");
  d.t_start = as_double(common, "t_start");
  d.t_mult = as_double(common, "t_mult");
  d.dt_max = as_double(common, "dt_max");
  const long long np = as_ll(common, "photons");
  if (np < 0) return err = "photons must be >= 0", false;
  d.n_photons = (uint64_t)np;
  d.seed = (uint32_t)(int)as_ll(common, "seed");

  const char* axes = "xyz";
  for (const auto& ch : spatial->children) {
    for (int a = 0; a < 3; ++a) {
      const std::string tag = std::string(1, axes[a]) + "_division";
      if (ch->name == tag) {
        Division v;
        v.start = as_double(ch.get(), (std::string(1, axes[a]) + "_start").c_str());
        v.end = as_double(ch.get(), (std::string(1, axes[a]) + "_end").c_str());
        v.cells = (uint32_t)as_ll(ch.get(), (std::string("n_") + axes[a] + "_cells").c_str());
        if (v.cells == 0) return err = tag + " with zero cells", false;
        d.div[a].push_back(v);
      }
    }
    if (ch->name == "region_map") {
      const uint32_t x = (uint32_t)as_ll(ch.get(), "x_div_ID");
      const uint32_t y = (uint32_t)as_ll(ch.get(), "y_div_ID");
      const uint32_t z = (uint32_t)as_ll(ch.get(), "z_div_ID");
      d.region_map[z * 1000000u + y * 1000u + x] = (uint32_t)as_ll(ch.get(), "region_ID");
    }
  }
  // boundary conditions (src/input.h:345-420): left/right/down/up/bottom/top
  const char* tags[6] = {"bc_left", "bc_right", "bc_down", "bc_up", "bc_bottom", "bc_top"};
  bool source_on = false;
  for (int f = 0; f < 6; ++f) {
    const std::string t = boundary->child_text(tags[f]);
    if (t == "REFLECT") d.bc[f] = BH_REFLECT;
    else if (t == "VACUUM") d.bc[f] = BH_VACUUM;
    else if (t == "SOURCE") {
      d.bc[f] = BH_SOURCE;
      source_on = true;
    } else {
      err = "ERROR: Boundary type not reconginzed (" + std::string(tags[f]) + "='" + t + "')";
      return false;
    }
  }
  if (source_on) d.T_source = as_double(boundary, "T_source");

  for (const bhx::Node* r : regions->children_named("region")) {
    Region g;
    g.id = (uint32_t)as_ll(r, "ID");
    g.cv = as_double(r, "CV");
    g.rho = as_double(r, "density");
    g.opA = as_double(r, "opacA");
    g.opB = as_double(r, "opacB");
    g.opC = as_double(r, "opacC");
    g.opS = as_double(r, "opacS");
    g.T_e = as_double(r, "initial_T_e");
    g.T_r = as_double(r, "initial_T_r");
    d.region_index[g.id] = (uint32_t)d.regions.size();
    d.regions.push_back(g);
  }
  if (d.regions.empty()) return err = "ERROR: No regions were specified.", false;
  const size_t n_div = d.div[0].size() * d.div[1].size() * d.div[2].size();
  if (n_div == 0) return err = "every axis needs at least one division", false;
  if (n_div != d.region_map.size())
    return err = "ERROR: Number of total divisions must match the number of unique region maps.",
           false;
  for (const auto& kv : d.region_map) {
    const uint32_t x = kv.first % 1000u, y = (kv.first / 1000u) % 1000u, z = kv.first / 1000000u;
    if (x >= d.div[0].size() || y >= d.div[1].size() || z >= d.div[2].size())
      return err = "region_map names a division that does not exist", false;
    if (!d.region_index.count(kv.second))
      return err = "region_map names region " + std::to_string(kv.second) + " which is not defined",
             false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Mesh: cells in proto-mesh order (src/proto_mesh.h:106-218) with material
// state (src/mesh.h:426-439) and step physics (src/mesh.h:237-423).
// ---------------------------------------------------------------------------
struct Mesh {
  uint32_t n = 0, G = 1;
  std::vector<double> nodes;    // [n*6]
  std::vector<int32_t> bc;      // [n*6]
  std::vector<uint32_t> next;   // [n*6]
  std::vector<uint32_t> region;  // index into Deck::regions
  std::vector<double> T_e, T_r_cell, T_s, cv, rho;
  // step physics
  std::vector<double> sigma_a, sigma_s, fleck;
  std::vector<double> emission_E, source_E, census_E;
  std::vector<int8_t> source_face;
  std::vector<double> T_r;  // diagnostic radiation temperature
  double total_photon_E = 0;   // this rank's photon energy
  double global_photon_E = 0;  // all ranks' photon energy (the source's E_tot)
  double replicated_factor = 1.0;

  double volume(uint32_t i) const {
    const double* nd = &nodes[(size_t)i * 6];
    return (nd[1] - nd[0]) * (nd[3] - nd[2]) * (nd[5] - nd[4]);
  }
  // Cell::get_face_area (src/cell.h:83-100)
  double face_area(uint32_t i, int face) const {
    const double* nd = &nodes[(size_t)i * 6];
    if (face == 0 || face == 1) return (nd[3] - nd[2]) * (nd[5] - nd[4]);
    if (face == 2 || face == 3) return (nd[1] - nd[0]) * (nd[5] - nd[4]);
    if (face == 4 || face == 5) return (nd[1] - nd[0]) * (nd[3] - nd[2]);
    return -1.0;
  }
};

void build_mesh(const Deck& d, uint32_t G, Mesh& m) {
  uint32_t ng[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a)
    for (const auto& v : d.div[a]) ng[a] += v.cells;
  const uint32_t n = ng[0] * ng[1] * ng[2];
  m.n = n;
  m.G = G;
  m.nodes.assign((size_t)n * 6, 0.0);
  m.bc.assign((size_t)n * 6, 0);
  m.next.assign((size_t)n * 6, 0);
  m.region.assign(n, 0);
  m.T_e.assign(n, 0);
  m.T_r_cell.assign(n, 0);
  m.T_s.assign(n, 0);
  m.cv.assign(n, 0);
  m.rho.assign(n, 0);
  m.source_face.assign(n, -1);
  uint32_t gc = 0, gk = 0;
  for (size_t izd = 0; izd < d.div[2].size(); ++izd) {
    const Division& dz = d.div[2][izd];
    const double hz = (dz.end - dz.start) / dz.cells;
    for (uint32_t k = 0; k < dz.cells; ++k, ++gk) {
      uint32_t gj = 0;
      for (size_t iyd = 0; iyd < d.div[1].size(); ++iyd) {
        const Division& dy = d.div[1][iyd];
        const double hy = (dy.end - dy.start) / dy.cells;
        for (uint32_t j = 0; j < dy.cells; ++j, ++gj) {
          uint32_t gi = 0;
          for (size_t ixd = 0; ixd < d.div[0].size(); ++ixd) {
            const Division& dx = d.div[0][ixd];
            const double hx = (dx.end - dx.start) / dx.cells;
            const uint32_t rid =
                d.region_map.at((uint32_t)izd * 1000000u + (uint32_t)iyd * 1000u + (uint32_t)ixd);
            const uint32_t ridx = d.region_index.at(rid);
            for (uint32_t i = 0; i < dx.cells; ++i, ++gi, ++gc) {
              double* nd = &m.nodes[(size_t)gc * 6];
              // the last cell of a division ends exactly where the next
              // division starts (src/proto_mesh.h:130-145)
              nd[0] = dx.start + i * hx;
              nd[1] = (i == dx.cells - 1 && ixd != d.div[0].size() - 1) ? d.div[0][ixd + 1].start
                                                                         : dx.start + (i + 1) * hx;
              nd[2] = dy.start + j * hy;
              nd[3] = (j == dy.cells - 1 && iyd != d.div[1].size() - 1) ? d.div[1][iyd + 1].start
                                                                         : dy.start + (j + 1) * hy;
              nd[4] = dz.start + k * hz;
              nd[5] = (k == dz.cells - 1 && izd != d.div[2].size() - 1) ? d.div[2][izd + 1].start
                                                                         : dz.start + (k + 1) * hz;
              m.region[gc] = ridx;
              int32_t* b = &m.bc[(size_t)gc * 6];
              uint32_t* nx = &m.next[(size_t)gc * 6];
              // neighbours (src/proto_mesh.h:156-202); a domain face keeps
              // the cell itself as its "next"
              const bool lo[3] = {gi > 0, gj > 0, gk > 0};
              const bool hi[3] = {gi < ng[0] - 1, gj < ng[1] - 1, gk < ng[2] - 1};
              const uint32_t stride[3] = {1u, ng[0], ng[0] * ng[1]};
              for (int a = 0; a < 3; ++a) {
                b[2 * a] = lo[a] ? BH_ELEMENT : d.bc[2 * a];
                nx[2 * a] = lo[a] ? gc - stride[a] : gc;
                b[2 * a + 1] = hi[a] ? BH_ELEMENT : d.bc[2 * a + 1];
                nx[2 * a + 1] = hi[a] ? gc + stride[a] : gc;
              }
              const Region& r = d.regions[ridx];
              m.cv[gc] = r.cv;
              m.rho[gc] = r.rho;
              m.T_e[gc] = r.T_e;
              m.T_r_cell[gc] = r.T_r;
              int face = -1;  // Cell::get_source_face (src/cell.h:69-76)
              for (int f = 0; f < 6; ++f)
                if (b[f] == BH_SOURCE) {
                  face = f;
                  break;
                }
              m.source_face[gc] = (int8_t)face;
              m.T_s[gc] = (face != -1) ? d.T_source : 0.0;
            }
          }
        }
      }
    }
  }
  m.sigma_a.assign((size_t)n * G, 0.0);
  m.sigma_s.assign((size_t)n * G, 0.0);
  m.fleck.assign(n, 0.0);
  m.emission_E.assign(n, 0.0);
  m.source_E.assign(n, 0.0);
  m.census_E.assign(n, 0.0);
  m.T_r.assign(n, 0.0);
}

// ---------------------------------------------------------------------------
// IMC_State (src/imc_state.h:42-134, 292-296)
// ---------------------------------------------------------------------------
struct State {
  double time = 0, dt = 0, t_stop = 0, t_mult = 0, dt_max = 0;
  uint32_t step = 1;
  double pre_mat_E = 0, emission_E = 0, source_E = 0, pre_census_E = 0;

  double next_dt() const {
    double nd = (dt * t_mult < dt_max) ? dt * t_mult : dt_max;
    if (time + nd > t_stop) nd = t_stop - time;
    return nd;
  }
  bool finished() const { return std::abs(time - t_stop) < 1.0e-8; }
  void advance() {
    time += dt;
    dt = next_dt();
    step++;
  }
};

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct bh_driver {
  std::string err;
  bh_driver_config cfg{};
  Deck deck;
  Mesh mesh;
  State st;
  bh_handle* eng = nullptr;
  std::vector<double> abs_E, track_E;
  double prev_census_E = 0.0;  // post-transport census energy of the last step
  uint64_t census_size = 0;
  enum Phase { kIdle, kPrepared, kTransported } phase = kIdle;
  bh_step_stats stats{};
  bh_resident_report rr{};      // device_physics: the last resident step
  bool resident_ready = false;  // device_physics: geometry and material uploaded
  uint64_t transported = 0;
  double prepare_ms = 0, device_ms = 0;
};

namespace {

int dfail(bh_driver* d, int code, const std::string& m) {
  if (d) d->err = m;
  return code;
}

int engine_fail(bh_driver* d, int rc) {
  return dfail(d, rc, std::string("engine: ") + bh_last_error(d->eng));
}

// Mesh::calculate_photon_energy (src/mesh.h:237-323) for this rank.
void calculate_photon_energy(bh_driver* d) {
  Mesh& m = d->mesh;
  State& s = d->st;
  const double dt = s.dt;
  const double rf = m.replicated_factor;
  const uint32_t n_ranks = d->cfg.n_ranks, rank = d->cfg.rank;
  double tot_census = 0, tot_emission = 0, tot_source = 0, pre_mat = 0;
  m.total_photon_E = 0;
  for (uint32_t i = 0; i < m.n; ++i) {
    const double vol = m.volume(i);
    const double cV = m.cv[i], T = m.T_e[i], Tr = m.T_r_cell[i], Ts = m.T_s[i], rho = m.rho[i];
    const Region& r = d->deck.regions[m.region[i]];
    const double op_a = r.absorption(T);
    const double op_s = r.opS;
    const double f = 1.0 / (1.0 + dt * op_a * kC * (4.0 * kA * std::pow(T, 3) / (cV * rho)));
    for (uint32_t g = 0; g < m.G; ++g) {  // Cell::set_op_a / set_op_s fill all groups
      m.sigma_a[(size_t)i * m.G + g] = op_a;
      m.sigma_s[(size_t)i * m.G + g] = op_s;
    }
    m.fleck[i] = f;
    m.emission_E[i] = rf * dt * vol * f * op_a * kA * kC * std::pow(T, 4);
    m.census_E[i] = (s.step > 1) ? 0.0 : rf * vol * kA * std::pow(Tr, 4);
    m.source_E[i] = rf * 0.25 * kA * kC * m.face_area(i, m.source_face[i]) * std::pow(Ts, 4) * dt;
    pre_mat += T * cV * vol * rho;
    tot_emission += m.emission_E[i];
    tot_census += m.census_E[i];
    tot_source += m.source_E[i];
    m.total_photon_E += m.source_E[i] + m.census_E[i] + m.emission_E[i];
  }
  // replicated small-energy reassignment (src/mesh.h:287-315).  The
  // all-reduce of the three step totals sums n_ranks equal values, taken in
  // rank order; a cell energy the all-rank source would give no photon of
  // its own goes whole to rank cell % n_ranks.
  const double local3 = tot_emission + tot_census + tot_source;
  double global_E = local3;
  for (uint32_t r = 1; r < n_ranks; ++r) global_E += local3;
  const uint32_t n_user = (uint32_t)d->deck.n_photons;  // n_user_photons is uint32 here
  auto reassign = [&](uint32_t i, uint32_t r, double& em, double& src, double& cen) {
    const bool mine = (i % n_ranks) == r;
    if (s.step == 1 && cen > 0.0 && int(n_user * (cen / global_E)) == 0)
      cen = mine ? cen / rf : 0.0;
    if (em > 0.0 && int(n_user * (em / global_E)) == 0) em = mine ? em / rf : 0.0;
    if (src > 0.0 && int(n_user * (src / global_E)) == 0) src = mine ? src / rf : 0.0;
  };
  // the source's E_tot (src/replicated_driver.h:57-59): the all-reduce of
  // every rank's total after its own reassignment, in rank order
  double all_ranks_E = 0.0;
  for (uint32_t r = 0; r < n_ranks && n_ranks > 1; ++r) {
    double t = 0.0;
    for (uint32_t i = 0; i < m.n; ++i) {
      double em = m.emission_E[i], src = m.source_E[i], cen = m.census_E[i];
      reassign(i, r, em, src, cen);
      t += src + cen + em;
    }
    all_ranks_E += t;
  }
  tot_census = tot_emission = tot_source = 0;
  m.total_photon_E = 0;
  for (uint32_t i = 0; i < m.n; ++i) {
    reassign(i, rank, m.emission_E[i], m.source_E[i], m.census_E[i]);
    tot_emission += m.emission_E[i];
    tot_census += m.census_E[i];
    tot_source += m.source_E[i];
    m.total_photon_E += m.source_E[i] + m.census_E[i] + m.emission_E[i];
  }
  // one rank: the all-reduce of its own total
  m.global_photon_E = (n_ranks == 1) ? m.total_photon_E : all_ranks_E;
  s.pre_mat_E = pre_mat;
  s.emission_E = tot_emission;
  s.source_E = tot_source;
  if (s.step == 1) s.pre_census_E = tot_census;
}

// Energy of the initial census exactly as the reference sums its photon list
// (get_photon_list_E, src/census_functions.h:40-46): n_c copies of
// E_c / n_c per cell, in cell order.
double initial_census_energy(const bh_driver* d, double total_E) {
  const Mesh& m = d->mesh;
  const uint64_t n_user = d->deck.n_photons;
  double sum = 0.0;
  for (uint32_t i = 0; i < m.n; ++i) {
    if (!(m.census_E[i] > 0.0)) continue;
    uint32_t n = (uint32_t)(int)((double)n_user * m.census_E[i] / total_E);
    if (n == 0) n = 1;
    const double E = m.census_E[i] / n;
    for (uint32_t p = 0; p < n; ++p) sum += E;
  }
  return sum;
}

}  // namespace

extern "C" {

int bh_driver_create(const char* xml_path, const bh_driver_config* cfg, bh_driver** out) {
  if (!out) return BH_ERR_ARG;
  *out = nullptr;
  bh_driver* d = new (std::nothrow) bh_driver();
  if (!d) return BH_ERR_ALLOC;
  *out = d;
  if (cfg) d->cfg = *cfg;
  if (d->cfg.n_groups == 0) d->cfg.n_groups = 1;
  if (d->cfg.n_ranks == 0) d->cfg.n_ranks = 1;
  if (d->cfg.rank >= d->cfg.n_ranks) return dfail(d, BH_ERR_ARG, "rank must be < n_ranks");
  if (!xml_path) return dfail(d, BH_ERR_ARG, "xml_path is NULL");
  std::string err;
  if (!parse_deck(xml_path, d->deck, err)) return dfail(d, BH_ERR_ARG, err);
  if (d->cfg.n_photons) d->deck.n_photons = d->cfg.n_photons;
  if (d->cfg.t_stop > 0) d->deck.t_stop = d->cfg.t_stop;
  if (d->deck.n_photons == 0) return dfail(d, BH_ERR_ARG, "photons must be > 0");
  if (d->deck.n_photons > 0xFFFFFFFFull)
    return dfail(d, BH_ERR_ARG, "photons must fit 32 bits (Mesh::calculate_photon_energy)");
  build_mesh(d->deck, d->cfg.n_groups, d->mesh);
  d->mesh.replicated_factor = 1.0 / (double)d->cfg.n_ranks;
  d->st.time = d->deck.t_start;
  d->st.dt = d->deck.dt;
  d->st.t_stop = d->deck.t_stop;
  d->st.t_mult = d->deck.t_mult;
  d->st.dt_max = d->deck.dt_max;
  d->abs_E.assign(d->mesh.n, 0.0);
  d->track_E.assign(d->mesh.n, 0.0);
  return BH_OK;
}

// The engine is created on first use so a driver can parse, mesh and compute
// step physics on a machine without a GPU.
static int ensure_engine(bh_driver* d) {
  if (d->eng) return BH_OK;
  const int rc = bh_create(d->cfg.device, &d->cfg.engine, &d->eng);
  if (rc != BH_OK) {
    d->err = std::string("engine: ") + (d->eng ? bh_last_error(d->eng) : "bh_create failed");
    if (d->eng) bh_destroy(d->eng);
    d->eng = nullptr;
  }
  return rc;
}

int bh_driver_calculate_physics(bh_driver* d) {
  if (!d) return BH_ERR_ARG;
  if (d->phase != bh_driver::kIdle) return dfail(d, BH_ERR_STATE, "step already prepared");
  if (d->st.finished()) return dfail(d, BH_ERR_STATE, "simulation finished");
  calculate_photon_energy(d);
  return BH_OK;
}

void bh_driver_destroy(bh_driver* d) {
  if (!d) return;
  if (d->eng) bh_destroy(d->eng);
  delete d;
}

const char* bh_driver_last_error(const bh_driver* d) { return d ? d->err.c_str() : "null driver"; }
int bh_driver_finished(const bh_driver* d) { return d ? (int)d->st.finished() : 1; }
uint32_t bh_driver_n_cells(const bh_driver* d) { return d ? d->mesh.n : 0; }
uint64_t bh_driver_n_photons(const bh_driver* d) { return d ? d->deck.n_photons : 0; }
uint32_t bh_driver_seed(const bh_driver* d) { return d ? d->deck.seed : 0; }
bh_handle* bh_driver_engine(bh_driver* d) {
  if (!d || ensure_engine(d) != BH_OK) return nullptr;
  return d->eng;
}

int bh_driver_prepare(bh_driver* d) {
  if (!d) return BH_ERR_ARG;
  if (d->phase != bh_driver::kIdle) return dfail(d, BH_ERR_STATE, "step already prepared");
  if (d->st.finished()) return dfail(d, BH_ERR_STATE, "simulation finished");
  const auto t0 = std::chrono::steady_clock::now();
  int rce = ensure_engine(d);
  if (rce != BH_OK) return rce;
  if (d->cfg.device_physics) {  // geometry and material once; physics on the device
    if (!d->resident_ready) {
      calculate_photon_energy(d);  // first opacities: the lattice test needs group-uniform ones
      Mesh& m = d->mesh;
      bh_cells c{m.n, m.G, m.nodes.data(), m.bc.data(), m.next.data(),
                 m.sigma_a.data(), m.sigma_s.data(), m.fleck.data()};
      int rc = bh_set_mesh(d->eng, &c);
      if (rc != BH_OK) return engine_fail(d, rc);
      std::vector<double> col[6];
      for (auto& v : col) v.resize(m.n);
      for (uint32_t i = 0; i < m.n; ++i) {
        const Region& r = d->deck.regions[m.region[i]];
        col[0][i] = r.opA;
        col[1][i] = r.opB;
        col[2][i] = r.opC;
        col[3][i] = r.opS;
        col[4][i] = m.cv[i];
        col[5][i] = m.rho[i];
      }
      bh_material mat{col[0].data(), col[1].data(), col[2].data(), col[3].data(),
                      col[4].data(), col[5].data(), m.T_e.data(),  m.T_r_cell.data(),
                      m.T_s.data(),  m.source_face.data()};
      rc = bh_set_material(d->eng, &mat, m.replicated_factor, d->cfg.n_ranks, d->cfg.rank);
      if (rc != BH_OK) return engine_fail(d, rc);
      d->resident_ready = true;
    }
    d->prepare_ms = ms_since(t0);
    d->phase = bh_driver::kPrepared;
    return BH_OK;
  }
  calculate_photon_energy(d);
  Mesh& m = d->mesh;
  bh_cells c{m.n, m.G, m.nodes.data(), m.bc.data(), m.next.data(),
             m.sigma_a.data(), m.sigma_s.data(), m.fleck.data()};
  int rc = bh_set_mesh(d->eng, &c);
  if (rc != BH_OK) return engine_fail(d, rc);
  rc = bh_set_source_energy(d->eng, m.emission_E.data(), m.source_E.data(),
                            d->st.step == 1 ? m.census_E.data() : nullptr, m.source_face.data());
  if (rc != BH_OK) return engine_fail(d, rc);
  d->prepare_ms = ms_since(t0);
  d->phase = bh_driver::kPrepared;
  return BH_OK;
}

int bh_driver_transport(bh_driver* d) {
  if (!d || !d->eng) return BH_ERR_ARG;
  if (d->phase != bh_driver::kPrepared) return dfail(d, BH_ERR_STATE, "call bh_driver_prepare first");
  const auto t0 = std::chrono::steady_clock::now();
  if (d->cfg.device_physics) {  // the whole step, physics included, on the device
    const int rc = bh_step_resident(d->eng, d->deck.n_photons, d->deck.seed, d->st.step,
                                    d->cfg.rank, d->st.dt, d->st.next_dt(), &d->rr);
    if (rc != BH_OK) return engine_fail(d, rc);
    d->stats = d->rr.stats;
    d->transported = d->rr.stats.n_transported;
    d->st.pre_census_E = (d->st.step == 1) ? d->rr.pre_census_E : d->prev_census_E;
    d->st.pre_mat_E = d->rr.pre_mat_E;
    d->st.emission_E = d->rr.emission_E;
    d->st.source_E = d->rr.source_E;
    d->mesh.global_photon_E = d->rr.total_E;
    d->device_ms = ms_since(t0);
    d->phase = bh_driver::kTransported;
    return BH_OK;
  }
  const double total_E = d->mesh.global_photon_E;
  if (!(total_E > 0.0)) return dfail(d, BH_ERR_STATE, "no photon energy this step");
  // pre-transport census energy (src/replicated_driver.h:61, 71)
  d->st.pre_census_E = (d->st.step == 1) ? initial_census_energy(d, total_E) : d->prev_census_E;
  if (d->cfg.engine.step_graph) {  // the whole device phase as one graph launch
    const int rc = bh_step_graph(d->eng, d->deck.n_photons, total_E, d->deck.seed, d->st.step,
                                 d->cfg.rank, d->st.dt, d->st.next_dt(), nullptr, nullptr,
                                 &d->stats);
    if (rc != BH_OK) return engine_fail(d, rc);
    d->transported = d->stats.n_transported;
  } else {
    uint64_t n_new = 0;
    int rc = bh_source(d->eng, d->deck.n_photons, total_E, d->deck.seed, d->st.step,
                       d->cfg.rank, d->st.dt, &n_new);
    if (rc != BH_OK) return engine_fail(d, rc);
    uint64_t n_cen = 0;
    bh_census_size(d->eng, &n_cen);
    d->transported = n_new + n_cen;
    rc = bh_transport(d->eng, d->st.next_dt(), nullptr, nullptr, &d->stats);
    if (rc != BH_OK) return engine_fail(d, rc);
  }
  d->device_ms = ms_since(t0);
  d->phase = bh_driver::kTransported;
  return BH_OK;
}

int bh_driver_finish(bh_driver* d, bh_driver_report* rep) {
  if (!d || !d->eng) return BH_ERR_ARG;
  if (d->phase != bh_driver::kTransported) return dfail(d, BH_ERR_STATE, "call bh_driver_transport first");
  const auto t0 = std::chrono::steady_clock::now();
  Mesh& m = d->mesh;
  State& s = d->st;
  double total_abs = 0, post_mat = 0;
  if (d->cfg.device_physics) {  // updated on the device by bh_step_resident
    total_abs = d->rr.absorbed_E;
    post_mat = d->rr.post_mat_E;
  } else {
  int rc = bh_get_tallies(d->eng, d->abs_E.data(), d->track_E.data());
  if (rc != BH_OK) return engine_fail(d, rc);
  // Mesh::update_temperature (src/mesh.h:327-423), one rank
  for (uint32_t i = 0; i < m.n; ++i) {
    const Region& r = d->deck.regions[m.region[i]];
    const double cV = r.cv, rho = r.rho, vol = m.volume(i), T = m.T_e[i];
    const double T_new = T + (d->abs_E[i] - m.emission_E[i]) / (cV * vol * rho);
    m.T_r[i] = std::pow(d->track_E[i] / (vol * s.dt * kA * kC), 0.25);
    m.T_e[i] = T_new;
    total_abs += d->abs_E[i];
    post_mat += T_new * cV * vol * rho;
  }
  }
  bh_driver_report R{};
  R.step = s.step;
  R.time = s.time;
  R.dt = s.dt;
  R.next_dt = s.next_dt();
  R.transported = d->transported;
  R.census_size = d->stats.n_census;
  R.n_segments = d->stats.n_segments;
  R.n_flushes = d->stats.n_flushes;
  R.n_replayed = d->stats.n_replayed;
  R.lattice = d->stats.lattice;
  R.tally = d->stats.tally;
  R.n_exit = d->stats.n_exit;
  R.n_killed = d->stats.n_killed;
  R.emission_E = s.emission_E;
  R.source_E = s.source_E;
  R.absorbed_E = total_abs;
  R.exit_E = d->stats.exit_E;
  R.pre_census_E = s.pre_census_E;
  R.post_census_E = d->stats.census_E;
  R.pre_mat_E = s.pre_mat_E;
  R.post_mat_E = post_mat;
  // IMC_State::print_conservation (src/imc_state.h:254-258)
  R.rad_conservation =
      (total_abs + R.post_census_E + R.exit_E) - (R.pre_census_E + R.emission_E + R.source_E);
  R.mat_conservation = post_mat - (R.pre_mat_E + total_abs - R.emission_E);
  R.source_ms = d->stats.source_ms;
  R.transport_ms = d->stats.transport_ms;
  R.census_ms = d->stats.census_ms;
  R.replay_ms = d->stats.replay_ms;
  R.n_allocs = d->stats.n_allocs;
  R.prepare_ms = d->prepare_ms;
  R.device_phase_ms = d->device_ms;
  d->prev_census_E = d->stats.census_E;
  d->census_size = d->stats.n_census;
  s.advance();
  R.finish_ms = ms_since(t0);
  if (rep) *rep = R;
  d->phase = bh_driver::kIdle;
  return BH_OK;
}

int bh_driver_step(bh_driver* d, bh_driver_report* rep) {
  int rc = bh_driver_prepare(d);
  if (rc == BH_OK) rc = bh_driver_transport(d);
  if (rc == BH_OK) rc = bh_driver_finish(d, rep);
  return rc;
}

int bh_driver_get_tallies(const bh_driver* d, double* abs_E, double* track_E) {
  if (!d) return BH_ERR_ARG;
  if (d->cfg.device_physics && d->eng) return bh_get_tallies(d->eng, abs_E, track_E);
  if (abs_E) std::memcpy(abs_E, d->abs_E.data(), d->mesh.n * sizeof(double));
  if (track_E) std::memcpy(track_E, d->track_E.data(), d->mesh.n * sizeof(double));
  return BH_OK;
}

int bh_driver_get_temperature(const bh_driver* d, double* T_e, double* T_r) {
  if (!d) return BH_ERR_ARG;
  if (d->cfg.device_physics && d->resident_ready) return bh_get_temperature(d->eng, T_e, T_r);
  if (T_e) std::memcpy(T_e, d->mesh.T_e.data(), d->mesh.n * sizeof(double));
  if (T_r) std::memcpy(T_r, d->mesh.T_r.data(), d->mesh.n * sizeof(double));
  return BH_OK;
}

int bh_driver_get_mesh(const bh_driver* d, double* nodes, int32_t* bc, uint32_t* next,
                       uint32_t* region) {
  if (!d) return BH_ERR_ARG;
  const size_t n = d->mesh.n;
  if (nodes) std::memcpy(nodes, d->mesh.nodes.data(), n * 6 * sizeof(double));
  if (bc) std::memcpy(bc, d->mesh.bc.data(), n * 6 * sizeof(int32_t));
  if (next) std::memcpy(next, d->mesh.next.data(), n * 6 * sizeof(uint32_t));
  if (region) std::memcpy(region, d->mesh.region.data(), n * sizeof(uint32_t));
  return BH_OK;
}

int bh_driver_get_physics(const bh_driver* d, double* sigma_a, double* sigma_s, double* fleck,
                          double* emission_E, double* source_E, double* census_E,
                          int8_t* source_face, double* total_E) {
  if (!d) return BH_ERR_ARG;
  const Mesh& m = d->mesh;
  const size_t n = m.n;
  if (d->cfg.device_physics && d->resident_ready) {
    const int rc = bh_get_physics(d->eng, sigma_a, sigma_s, fleck, emission_E, source_E);
    if (rc != BH_OK) return rc;
    if (census_E) std::memset(census_E, 0, n * sizeof(double));
    if (source_face) std::memcpy(source_face, m.source_face.data(), n);
    if (total_E) *total_E = d->rr.total_E;  // E_tot of the last resident step
    return BH_OK;
  }
  if (sigma_a) std::memcpy(sigma_a, m.sigma_a.data(), n * m.G * sizeof(double));
  if (sigma_s) std::memcpy(sigma_s, m.sigma_s.data(), n * m.G * sizeof(double));
  if (fleck) std::memcpy(fleck, m.fleck.data(), n * sizeof(double));
  if (emission_E) std::memcpy(emission_E, m.emission_E.data(), n * sizeof(double));
  if (source_E) std::memcpy(source_E, m.source_E.data(), n * sizeof(double));
  if (census_E) std::memcpy(census_E, m.census_E.data(), n * sizeof(double));
  if (source_face) std::memcpy(source_face, m.source_face.data(), n);
  if (total_E) *total_E = m.global_photon_E;
  return BH_OK;
}

}  // extern "C"
