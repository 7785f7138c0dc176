// rewl_run -- a small host program over rewl_host.hpp, the counterpart of the
// reference's main() (main.cpp:18-418) without its file tree: it runs one REWL
// simulation of a model chosen at compile time (-DREWL_MODEL_AT for
// Ashkin-Teller, Ising otherwise, as model_hamiltonians.hpp chooses by macro)
// and prints the normalised microcanonical ln g/N per energy bin.
//
//   rewl_run L windows replicas final_increment [average_mode] [observables] [densities] [anneal]
//
// With observables = 1 it also prints the per-bin means of the model's
// observables (E/N, then one column per observable); with observables = 2
// the Fourier correlator G(qmin) columns as well.  With densities = 1
// (Ashkin-Teller build, observables >= 1) it prints the sigma-tau density plot
// of every bin, 1024 values a line (the reference's AT_DENSITIES output).
// With anneal = 1 it first anneals 64 chains on the GPU (20 temperature steps,
// blocks of 200 sweeps; the SIMULATED_ANNEALING start of main.cpp:56-76),
// takes the lowest energy met as the bottom of the range and starts every
// walker from that chain's final spins.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "rewl_host.hpp"

#ifdef REWL_MODEL_AT
template <typename T>
using Hamiltonian_t = rewl_host::Ashkin_Teller2d<T>;
#else
template <typename T>
using Hamiltonian_t = rewl_host::Ising2d<T>;
#endif
using System_Parameters = Hamiltonian_t<double>::Parameters;

int main(int argc, char** argv) {
    System_Parameters sp;
    sp.L = argc > 1 ? std::atoi(argv[1]) : 8;
    rewl_host::REWL_Parameters rp;
    rp.num_windows = argc > 2 ? std::atoi(argv[2]) : 2;
    rp.replicas_per_window = argc > 3 ? std::atoi(argv[3]) : 2;
    rp.final_increment = argc > 4 ? std::atof(argv[4]) : 1e-4;
    rp.average_mode = argc > 5 ? std::atoi(argv[5]) : 0;
    rp.observables = argc > 6 ? std::atoi(argv[6]) : 0;
    rp.densities = argc > 7 && std::atoi(argv[7]) != 0;
    const bool anneal = argc > 8 && std::atoi(argv[8]) != 0;
    try {
        double e_min = sp.energy_min();
        rewl_host::Annealed start;
        if (anneal) {
            rewl_host::SA_Parameters sa;
            sa.num_iterations = 20;
            sa.block_size = 200;
            sa.chains = 64;
            start = rewl_host::simulated_annealing<Hamiltonian_t<double>>(sp, sa, rp.seed);
            e_min = start.ground_state_energy;
            std::printf("# annealed ground state %.15e\n", e_min);
        }
        rewl_host::REWL_simulation<Hamiltonian_t<double>> sim(sp, rp, e_min, sp.energy_max(), sp.energy_bin_size());
        if (anneal) sim.import_DoFs(start.spins);
        const bool done = sim.simulate(1000000);
        std::vector<double> E, lg;
        sim.results(E, lg);
        const rewl_stats st = sim.stats();
        std::printf("# converged %d sweeps %llu flat_events %llu exchanges %llu\n", done ? 1 : 0,
                    (unsigned long long)st.sweeps, (unsigned long long)st.flat_events,
                    (unsigned long long)st.exchanges_accepted);
        std::printf("# E/N, ln g/N\n");
        for (size_t i = 0; i < E.size(); ++i) std::printf("%.15e,%.15e\n", E[i] / sp.N(), lg[i] / sp.N());
        if (rp.observables) {
            std::vector<double> Eo, means;
            int nobs = 0;
            sim.observables(Eo, means, nobs);
            std::printf("# observables: E/N, then %d per-bin means\n", nobs);
            for (size_t i = 0; i < Eo.size(); ++i) {
                if (std::isnan(means[i * nobs])) continue;
                std::printf("%.15e", Eo[i] / sp.N());
                for (int o = 0; o < nobs; ++o) std::printf(",%.15e", means[i * nobs + o]);
                std::printf("\n");
            }
        }
        if (rp.densities) {
            std::vector<std::vector<double>> plots;
            sim.density_plots(plots);
            std::printf("# density plots: %zu bins x %d cells\n", plots.size(), REWL_DENSITY_CELLS);
            for (const auto& plot : plots) {
                for (size_t i = 0; i < plot.size(); ++i) std::printf(i ? ",%.15e" : "%.15e", plot[i]);
                std::printf("\n");
            }
        }
        return done ? 0 : 3;
    } catch (const rewl_host::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
}
