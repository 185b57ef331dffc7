// Host-side FP64 set-up: constants, small tables and GPU table-builder factors.
#pragma once

#include <vector>

#include "tempo_const.hpp"

namespace tempo {

// literal decade bins of the reference (src/module_mp_tempo_params.F90:194-263)
struct Bins {
  double r_c[37], r_i[64], r_r[37], r_s[37], r_g[37], n0r_exp[37], n0g_exp[37], nt_i[55], nt_in[55];
};

struct HostBins {
  Bins lit;
  double dc[kNbins], dtc[kNbins], di[kNbins], dti[kNbins], dr[kNbins], dtr[kNbins];
  double ds[kNbins], dts[kNbins], dg[kNbins], dtg[kNbins], t_nc[kNbins];
  double xm0g;
};

// Fortran-order FP64 arrays of the init-time tables
struct SmallTables {
  std::vector<double> efrw, efsw;            // (100,100)
  std::vector<double> tnc_wev;               // (100,37,100)
  std::vector<double> tps_iaus, tni_iaus, tpi_ide;  // (64,55)
};

// Factor arrays consumed by the GPU table builder (layouts documented in
// tempo_tables.hip).
struct TableFactors {
  std::vector<double> n_r;            // [m][k][n2]
  std::vector<double> qg_a, qg_ng;    // [q][n3][n2][n], [n3][j][i][n]
  std::vector<double> qs_a, qs_ns;    // [q][n2][n], [j][i][n]
  std::vector<double> fz_massr, fz_massc, fz_prob_r, fz_prob_c, fz_nc, fz_t_nc;
  std::vector<double> r_r, r_c;
  double fz_xm0g = 0.0;
};

Bins make_bins();
void build_const(Const<double>& c, HostBins& hb, bool hail, bool cube_as_product = false);
void build_small_tables(const Const<double>& c, const HostBins& hb, SmallTables& t);
// constants of the melting-layer reflectivity
void build_radar(const Const<double>& c, Radar& r);
void build_table_factors(TableFactors& f);
void snow_m2_m3(const Const<double>& c, double rs, double tc, double& smob, double& smoc);
double gamma_p_host(double a, double x);

}  // namespace tempo
