// remc_sim -- command-line REMC docking simulation (the reference's main(),
// "simulation engine/src/main.cpp:1575-1851"), running every replica on one
// device through the C ABI.
//
//   remc_sim -f config [-m mcsteps] [-e resteps] [-r replicas] [-b box]
//            [-t0 tmin] [-t1 tmax] [--seed s] [--outdir dir] ...
//
// The force-field data (AminoAcids, pair_potentials_MJ1996) is read from
// --data, else from the data/ directory beside this executable.
#include <climits>
#include <cstdio>
#include <ctime>
#include <string>
#include <unistd.h>

#include "remc_facade.h"

using namespace remc_facade;

static std::string exe_dir()
{
    char buf[PATH_MAX];
    ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
    if (n <= 0) return ".";
    buf[n] = 0;
    std::string s(buf);
    size_t k = s.find_last_of('/');
    return k == std::string::npos ? "." : s.substr(0, k);
}

int main(int argc, char **argv)
{
    argdata p;
    p.pid = (int)getpid();
    if (!getArgs(&p, argc, argv)) return 1;
    if (p.help) return 0;
    Checkpoint ck;
    if (p.resume) {
        // the checkpoint names the configuration and holds the run's parameters
        if (!loadSimulation(p.resumeFile, &p, &ck)) return 1;
        std::printf("The simulation will resume from the checkpoint file: %s\nOriginal input: %s\n",
                    p.resumeFile.c_str(), p.file.c_str());
    }
    if (!p.inputFile && !p.check) {
        std::printf("! no configuration file given (-f); see -h\n");
        return 1;
    }
    const std::string data = p.dataDir.empty() ? exe_dir() + "/data" : p.dataDir;
    AminoAcids aas;
    std::printf("Loading amino acid data. %s/AminoAcids\n", data.c_str());
    if (!aas.loadAminoAcidData((data + "/AminoAcids").c_str())) return 1;
    std::printf("Loading pair lookup table. %s/pair_potentials_MJ1996\n", data.c_str());
    if (!aas.loadLJPotentialData((data + "/pair_potentials_MJ1996").c_str())) return 1;
    if (p.check) return checkConformations(aas, p); // main.cpp:1652-1707 exits after the check

    Replica initial;
    initial.setAminoAcidData(aas);
    initial.setBoundingValue(150.0f);
    p.nonCrowders = 0;
    if (!loadArgsFromFile(&p, &initial)) return 1;
    const int nonCrowders = p.nonCrowders;
    if (p.resume && !loadSimulation(p.resumeFile, &p, &ck)) return 1; // its parameters win over the file's
    // second pass: command-line values override the file (main.cpp:1686)
    if (!getArgs(&p, argc, argv)) return 1;
    p.nonCrowders = nonCrowders;
    initial.setBoundingValue(p.bound);
    if (!p.seedGiven) p.seed = ((uint64_t)time(nullptr) << 20) ^ (uint64_t)p.pid;

    std::printf("Loaded: %d residues in %d molecules:\n", initial.residueCount(), (int)initial.molecules.size());
    for (size_t i = 0; i < initial.molecules.size(); ++i)
        std::printf("%2d: %3d %s\n", (int)i, (int)initial.molecules[i].residueCount(),
                    initial.molecules[i].filename.c_str());
    std::printf("%ld residue interaction pairs.\n", initial.countpairs());
    if (!p.savePdb.empty()) initial.saveAsSinglePDB(p.savePdb.c_str());
    if (initial.nonCrowderCount < (int)initial.molecules.size())
        std::printf("-!- Crowding modelled using: u(r) = (6/r)^ 6.\n");
    if (p.dryRun) {
        std::printf("Finished.\n");
        return 0;
    }

    const double e = checkedInitialEnergy(&initial, p.device);
    if (e != e) return 1;
    std::printf("RNG seed %llu\n", (unsigned long long)p.seed);
    std::printf("Output files will be prefixed by %s_%d\n", p.prependageString.c_str(), p.pid);
    std::vector<Replica> finals;
    int rc = REMCSimulation(&initial, &p, p.finalPdb.empty() ? nullptr : &finals, p.resume ? &ck : nullptr);
    if (rc != 0) return rc;
    if (!p.finalPdb.empty()) {
        size_t lo = 0;
        for (size_t i = 1; i < finals.size(); ++i)
            if (finals[i].temperature < finals[lo].temperature) lo = i;
        finals[lo].saveAsSinglePDB(p.finalPdb.c_str());
    }
    std::printf("Simulation done\nFinished.\n");
    return 0;
}
