// remc_pdbgen -- PDB files of sampled bound conformations ("post processor/
// pdbgensrc/pdbgen.cpp:18-235"): reads a run's fileindex and
// boundconformations files, rebuilds every record at the chosen temperature
// from its printed rotations and centres, and writes it as a single PDB file
// (and, optionally, as one MODEL of a multi-model trajectory), with its DRMS
// against the loaded (input) conformation.
//
//   remc_pdbgen <prefix>_fileindex <boxdim> [0/1: reverse-order DRMS] [1: include crowders]
//               [--temperature T] [--after N] [--outdir d] [--trajectory f] [--data d]
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

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
    std::printf("usage: remc_pdbgen <indexfile> <boxdim> [0/1: reverse drms cal] [1: outputcrowders crowders] "
                "[--temperature T] [--after N] [--outdir d] [--trajectory f] [--data d]\n");
    std::vector<std::string> pos;
    float target = 300.0f;
    long long after = 10000000;
    std::string outdir = "pdbgen_results", traj, data = exe_dir() + "/data";
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a.rfind("--", 0) == 0) {
            if (i + 1 >= argc) {
                std::printf("missing value for %s\n", a.c_str());
                return 1;
            }
            const char *v = argv[++i];
            if (a == "--temperature") target = (float)std::atof(v);
            else if (a == "--after") after = std::atoll(v);
            else if (a == "--outdir") outdir = v;
            else if (a == "--trajectory") traj = v;
            else if (a == "--data") data = v;
            else {
                std::printf("unknown option %s\n", a.c_str());
                return 1;
            }
        } else {
            pos.push_back(a);
        }
    }
    if (pos.size() < 2) return 1;
    const std::string index = pos[0];
    const float boundary = (float)std::atof(pos[1].c_str());
    const bool reverse = pos.size() > 2 && pos[2][0] == '1';
    const bool crowders = pos.size() > 3 && pos[3][0] == '1';
    std::printf("File: %s\nBounding box size: %gA**3\n", index.c_str(), (double)boundary);
    const size_t at = index.rfind("fileindex");
    if (at == std::string::npos) {
        std::printf("You must provide and index file as the input\n");
        return 1;
    }
    AminoAcids aas;
    if (!aas.loadAminoAcidData((data + "/AminoAcids").c_str()) ||
        !aas.loadLJPotentialData((data + "/pair_potentials_MJ1996").c_str()))
        return 1;

    // molecules of the index: non-crowders, and crowders when asked for
    Replica initial;
    initial.setAminoAcidData(aas);
    std::vector<char> is_crowder; // per loaded molecule
    std::vector<int> sim_index;   // its index in the simulation
    {
        FILE *f = std::fopen(index.c_str(), "r");
        if (!f) {
            std::printf("Failed to open file: %s\n", index.c_str());
            return 1;
        }
        char line[1024], name[512];
        while (std::fgets(line, sizeof line, f)) {
            int idx;
            char c;
            if (std::sscanf(line, "%d %511s %c", &idx, name, &c) != 3) continue;
            if (c == 'N' || (c == 'Y' && crowders)) {
                std::printf("load : %s\n", name);
                if (initial.loadMolecule(name) < 0) return 1;
                is_crowder.push_back(c == 'Y');
                sim_index.push_back(idx);
            }
        }
        std::fclose(f);
    }
    if (initial.molecules.empty()) {
        std::printf("no molecules in %s\n", index.c_str());
        return 1;
    }
    std::printf("finished loading\n");
    const std::string base = index.substr(0, at);
    const std::string conf = base + "boundconformations";
    FILE *df = std::fopen(conf.c_str(), "r");
    if (!df) {
        std::printf("Failed to open file: %s\n", conf.c_str());
        return 1;
    }
    std::printf("loading data: %s\n", conf.c_str());
    mkdir(outdir.c_str(), 0755);
    FILE *tf = traj.empty() ? nullptr : std::fopen(traj.c_str(), "w");
    FILE *drmsf = std::fopen((base + "drms.dat").c_str(), "w");
    if ((!traj.empty() && !tf) || !drmsf) {
        std::printf("cannot create the output files\n");
        return 1;
    }

    // records: "iteration; complex (replica); temperature" then one line per
    // simulation molecule "crowder index: w x y z cx cy cz"
    struct Pose {
        int crowder, index;
        double q[4];
        double c[3]; // read in double and rounded to float, as atof into a float does
    };
    char line[1024];
    std::vector<std::string> lines;
    while (std::fgets(line, sizeof line, df)) lines.push_back(line);
    std::fclose(df);
    int instance_no = 0;
    for (size_t li = 0; li < lines.size();) {
        int iteration;
        float energy, energy_c, temperature;
        if (std::sscanf(lines[li].c_str(), "%d; %f (%f); %f", &iteration, &energy, &energy_c, &temperature) != 4) {
            ++li;
            continue;
        }
        std::vector<Pose> poses;
        for (++li; li < lines.size(); ++li) {
            Pose p;
            if (std::sscanf(lines[li].c_str(), "%d %d: %lf %lf %lf %lf %lf %lf %lf", &p.crowder, &p.index, &p.q[0], &p.q[1],
                            &p.q[2], &p.q[3], &p.c[0], &p.c[1], &p.c[2]) != 9)
                break;
            poses.push_back(p);
        }
        if (!(std::fabs(temperature - target) < 1.0f && iteration > after)) continue;
        Replica inst = initial;
        inst.temperature = temperature;
        inst.potential = energy;
        bool complete = true;
        for (size_t i = 0; i < inst.molecules.size(); ++i) {
            const Pose *p = nullptr;
            for (const Pose &c : poses)
                if (c.index == sim_index[i]) p = &c;
            if (!p) {
                complete = false;
                break;
            }
            Quaternion q(p->q[0], p->q[1], p->q[2], p->q[3]);
            q.normalize();
            inst.molecules[i].setRotation(q);
            Vector3f v((float)p->c[0], (float)p->c[1], (float)p->c[2]);
            if (!is_crowder[i] && boundary > 0.0f) {
                // the periodic image closest to the first molecule
                const Vector3f &p0 = inst.molecules[0].position;
                const Vector3f d(v.x - p0.x, v.y - p0.y, v.z - p0.z);
                v.x -= boundary * std::round(d.x / boundary);
                v.y -= boundary * std::round(d.y / boundary);
                v.z -= boundary * std::round(d.z / boundary);
            }
            inst.molecules[i].setPosition(v);
        }
        if (!complete) continue;
        ++instance_no;
        char name[512];
        std::snprintf(name, sizeof name, "%s/sample_%d_%0.1fK_%5.2f.pdb", outdir.c_str(), instance_no,
                      (double)temperature, (double)energy);
        inst.saveAsSinglePDB(name);
        const double d = DRMS(inst, initial, reverse);
        std::fprintf(drmsf, "%f %f\n", (double)energy, d);
        if (tf) inst.writeToTrajectory(tf, instance_no, d);
        std::printf("Saved frame %d: %s\n", instance_no, name);
    }
    if (tf) std::fclose(tf);
    std::fclose(drmsf);
    std::printf("Exiting...\n");
    return 0;
}
