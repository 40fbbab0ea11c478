// Drop-in driver for the reference's GMRES/GMRES.cpp main (321-651): the same
// 12 positional inputs, the same GMRES sweeps and the same files under
// result/result_<Qchoice>_<nChebNodes>_<treeAdaptivity>_<int kappa>_<m>/:
//   Phi_<restart>_<m>   one "(re,im)" per line, 6 significant digits
//   gridPointsX, gridPointsY, solutionR, solutionI, leftRightBoundary,
//   bottomTopBoundary
// plus full-precision sidecars Phi_<restart>_<m>.f64 (binary complex128).
// The tree is the reference's adaptive one (createAdaptiveTree +
// makeLevelRestriction_Tree, GMRES.cpp:165-166, leaves resolving the contrast to
// 10^-treeAdaptivity). Extra optional arguments after the 12:
//   argv[13] incidence angle, argv[14] bump seed (Qchoice 6), argv[15] #bumps,
//   argv[16] tree: 1 adaptive (default), 0 uniform of depth nLevelsUniform (the
//   reference's commented-out createUniformTree, GMRES.cpp:164).
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "lsg_dfmm.hpp"

using namespace lsg;

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void write_phi(const std::string& dir, const std::string& name, const Vec& Phi) {
  std::ofstream f(dir + "/" + name);
  for (const auto& v : Phi) f << v << std::endl;
  std::ofstream g(dir + "/" + name + ".f64", std::ios::binary);
  g.write(reinterpret_cast<const char*>(Phi.data()), std::streamsize(Phi.size() * sizeof(Phi[0])));
}

int main(int argc, char* argv[]) {
  if (argc < 13) {
    std::cerr << "usage: " << argv[0]
              << " nCones_LFR nChebNodes treeAdaptivity L kappa yes2DFMM degreeOfBases TOL_POW m restart "
                 "nLevelsUniform Qchoice [angle seed nbumps tree]\n";
    return 1;
  }
  inputsToDFMM in;
  in.nCones_LFR = std::atoi(argv[1]);
  in.nChebNodes = std::atoi(argv[2]);
  in.treeAdaptivity = std::atoi(argv[3]);
  in.L = std::atof(argv[4]);
  in.kappa = std::atof(argv[5]);
  in.yes2DFMM = std::atoi(argv[6]);
  in.degreeOfBases = std::atoi(argv[7]);
  in.TOL_POW = std::atoi(argv[8]);
  const int m = std::atoi(argv[9]);
  const int restart = std::atoi(argv[10]);
  in.nLevelsUniform = std::atoi(argv[11]);
  in.Qchoice = std::atoi(argv[12]);
  if (argc > 13) in.inc_angle = std::atof(argv[13]);
  if (argc > 16) in.tree_mode = std::atoi(argv[16]);
  if (in.Qchoice == 6) {
    unsigned long long seed = argc > 14 ? std::strtoull(argv[14], nullptr, 10) : 0;
    int nb = argc > 15 ? std::atoi(argv[15]) : 5;
    lsg_bumps_from_seed(seed, nb, -0.35, 0.35, 0.03, 0.10, 0.5, 1.5, &in.contrast);
  }
  const double PI = 3.1415926535897932384;
  std::cout << "Wavenumber:\t\t" << in.kappa << std::endl;
  std::cout << "Wavelength:\t\t" << 2 * PI / in.kappa << std::endl;
  std::cout << "no. of full cycles:\t" << in.L * in.kappa / 2 * PI << std::endl;
  try {
    double t0 = now();
    DFMM S(in);
    S.FMMTreeInitialisation();
    std::cout << std::endl << "Time taken to assemble: " << now() - t0 << std::endl;
    const int N = S.N;
    std::cout << "nLevels: " << S.nLevels << std::endl;
    std::cout << "level_LFR: " << S.level_LFR << std::endl;
    std::cout << "Number of particles: " << N << std::endl;
    Vec b = S.rhs();
    mkdir("result", 0775);
    const std::string dir = "result/result_" + std::to_string(in.Qchoice) + "_" + std::to_string(in.nChebNodes) +
                            "_" + std::to_string(in.treeAdaptivity) + "_" + std::to_string(int(in.kappa)) + "_" +
                            std::to_string(m);
    mkdir(dir.c_str(), 0775);
    Vec Phi, APhi;
    // the reference's sweep (GMRES/GMRES.cpp:383-535), including its file
    // name for the (2,500) run
    const int sweep[10][2] = {{restart, 10}, {restart, 50}, {restart, 100}, {restart, 200}, {restart, 500},
                              {2, 10},       {2, 50},       {2, 100},       {2, 200},       {2, 500}};
    for (int s = 0; s < 10; ++s) {
      const int r = sweep[s][0], mm = sweep[s][1];
      t0 = now();
      RestartedGMRES<DFMM>(r, &S, mm, in, b, Phi);
      const double tg = now() - t0;
      std::string name = "Phi_" + std::to_string(s < 5 ? 1 : 2) + "_" + std::to_string(s == 9 ? 50 : mm);
      write_phi(dir, name, Phi);
      std::cout << std::endl << "Time taken by GMRES Solver " << r << "," << mm << ": " << tg << std::endl;
      S.MatVecProduct(Phi, APhi);
      double rn = 0, bn = 0;
      for (int i = 0; i < N; ++i) { rn += std::norm(b[i] - APhi[i]); bn += std::norm(b[i]); }
      std::cout << "Err in GMRES: " << std::sqrt(rn / bn) << std::endl << std::endl;
    }
    // scattered field from Phi with the contrast-free operator (GMRES.cpp:580-646)
    DFMM S2(in);
    S2.findPhi = false;
    S2.FMMTreeInitialisation();
    Vec Uinc = S2.incident(), Uscat;
    t0 = now();
    S2.MatVecProduct(Phi, Uscat);
    std::cout << std::endl << "Time taken for MatVec product: " << now() - t0 << std::endl;
    std::ofstream fx(dir + "/gridPointsX"), fy(dir + "/gridPointsY"), fr(dir + "/solutionR"), fi(dir + "/solutionI");
    for (int i = 0; i < N; ++i) {
      fx << S.gridPoints[i].x << std::endl;
      fy << S.gridPoints[i].y << std::endl;
      const std::complex<double> ut = Uinc[i] + Uscat[i];
      fr << ut.real() << std::endl;
      fi << ut.imag() << std::endl;
    }
    std::vector<pts2D> lr, bt;
    S2.boundaries(lr, bt);
    std::ofstream f5(dir + "/leftRightBoundary"), f6(dir + "/bottomTopBoundary");
    for (auto& p : lr) f5 << p.x << "\t" << p.y << std::endl;
    for (auto& p : bt) f6 << p.x << "\t" << p.y << std::endl;
  } catch (const Error& e) {
    std::cerr << "error " << e.code << ": " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
