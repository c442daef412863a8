// remc_facade.cpp -- the reference's host-side REMC API over the C ABI.
//
// Reference paths below are relative to "simulation engine/src/".  Rigid-body
// arithmetic reuses remc_math.h (the same quaternion product and rotation the
// kernels use), so a conformation built here and one moved on the device are
// computed the same way.
#include "remc_facade.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <sys/stat.h>
#include <unistd.h>

#include "../remc_math.h"

namespace remc_facade {

// ------------------------------------------------------------ vectors ----

void Vector3double::normalizeInPlace()
{
    // vector3f.cpp:36-43
    double m = std::sqrt(x * x + y * y + z * z);
    x /= m;
    y /= m;
    z /= m;
}

Quaternion Quaternion::normalize()
{
    // Quaternion.cpp:19-27
    double m = std::sqrt(w * w + x * x + y * y + z * z);
    w /= m;
    x /= m;
    y /= m;
    z /= m;
    return *this;
}

Quaternion Quaternion::operator*(const Quaternion &q) const
{
    const double a[4] = {w, x, y, z}, b[4] = {q.w, q.x, q.y, q.z};
    double o[4];
    remc_qmul(a, b, o);
    return Quaternion(o[0], o[1], o[2], o[3]);
}

Vector3f Quaternion::rotateVector(Vector3f v) const
{
    const double q[4] = {w, x, y, z};
    double m[9];
    remc_qcoeffs(q, m);
    Vector3f o;
    remc_qrotate(m, v.x, v.y, v.z, &o.x, &o.y, &o.z);
    return o;
}

// ------------------------------------------------------- amino acids ----

static std::string rstrip_nl(const std::string &s)
{
    size_t e = s.size();
    while (e > 0 && (s[e - 1] == '\n' || s[e - 1] == '\r')) --e;
    return s.substr(0, e);
}

static bool read_lines(const char *path, std::vector<std::string> &out)
{
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char buf[1024];
    while (std::fgets(buf, sizeof buf, f)) out.push_back(rstrip_nl(buf));
    std::fclose(f);
    return true;
}

bool AminoAcids::loadAminoAcidData(const char *filename)
{
    // AminoAcids.cpp:25-71: blocks of five lines (name, 3-letter name,
    // symbol, vdW radius, charge); '#' and blank lines skipped.
    std::vector<std::string> lines;
    if (!read_lines(filename, lines)) {
        std::fprintf(stderr, "Failed to open amino-acid data %s\n", filename);
        return false;
    }
    count = 0;
    for (size_t i = 0; i < lines.size();) {
        const std::string &ln = lines[i];
        if (ln.empty() || ln[0] == '#') {
            ++i;
            continue;
        }
        if (i + 4 >= lines.size() || count >= REMC_NTYPES) {
            std::fprintf(stderr, "Malformed amino-acid data %s at line %zu\n", filename, i + 1);
            return false;
        }
        AminoAcid &a = data[count++];
        a.name = ln.substr(0, 19);
        a.sname = lines[i + 1].substr(0, 3);
        a.symbol = lines[i + 2].empty() ? ' ' : lines[i + 2][0];
        a.vanderWaalRadius = (float)std::strtod(lines[i + 3].c_str(), nullptr);
        a.electrostaticCharge = (float)std::strtod(lines[i + 4].c_str(), nullptr);
        i += 5;
    }
    if (count != REMC_NTYPES) {
        std::fprintf(stderr, "Expected %d amino acids in %s, found %d\n", REMC_NTYPES, filename, count);
        return false;
    }
    return true;
}

int AminoAcids::getAminoAcidIndex(const char *name) const
{
    // AminoAcids.cpp:146-195: 3-letter name (capitalised), 1-letter symbol or
    // full name, searched from the end of the table.
    size_t n = std::strlen(name);
    if (n == 0) return -1;
    std::string key(name);
    if (n == 3) {
        key[0] = (char)std::toupper((unsigned char)key[0]);
        for (size_t k = 1; k < 3; ++k) key[k] = (char)std::tolower((unsigned char)key[k]);
        for (int i = count - 1; i >= 0; --i)
            if (data[i].sname == key) return i;
    } else if (n == 1) {
        char s = (char)std::toupper((unsigned char)key[0]);
        for (int i = count - 1; i >= 0; --i)
            if (data[i].symbol == s) return i;
    } else {
        key[0] = (char)std::toupper((unsigned char)key[0]);
        for (int i = count - 1; i >= 0; --i) {
            std::string nm = data[i].name;
            while (!nm.empty() && nm.back() == ' ') nm.pop_back();
            if (nm == key) return i;
        }
    }
    return -1;
}

bool AminoAcids::loadLJPotentialData(const char *filename)
{
    // AminoAcids.cpp:73-139: a label row, then the upper triangle row by row;
    // stored symmetric in the amino-acid table's order.
    std::vector<std::string> lines;
    if (!read_lines(filename, lines)) {
        std::fprintf(stderr, "Failed to open pair-potential table %s\n", filename);
        return false;
    }
    int labels[REMC_NTYPES];
    bool have_labels = false;
    int row = 0;
    for (const std::string &ln : lines) {
        if (ln.empty() || ln[0] == '#') continue;
        std::vector<std::string> toks;
        {
            char *buf = strdup(ln.c_str());
            for (char *t = std::strtok(buf, " \t"); t; t = std::strtok(nullptr, " \t")) toks.push_back(t);
            std::free(buf);
        }
        if (toks.empty()) continue;
        if (!have_labels) {
            if ((int)toks.size() != REMC_NTYPES) {
                std::fprintf(stderr, "Pair table %s: label row has %zu entries\n", filename, toks.size());
                return false;
            }
            for (int k = 0; k < REMC_NTYPES; ++k) {
                labels[k] = getAminoAcidIndex(toks[k].c_str());
                if (labels[k] < 0) {
                    std::fprintf(stderr, "Pair table %s: unknown residue %s\n", filename, toks[k].c_str());
                    return false;
                }
            }
            have_labels = true;
            continue;
        }
        if (row >= REMC_NTYPES || (int)toks.size() < 1 + REMC_NTYPES - row) {
            std::fprintf(stderr, "Pair table %s: malformed row %d\n", filename, row);
            return false;
        }
        for (int j = row; j < REMC_NTYPES; ++j) {
            float v = (float)std::strtod(toks[1 + j - row].c_str(), nullptr);
            LJpotentials[labels[row]][labels[j]] = v;
            LJpotentials[labels[j]][labels[row]] = v;
        }
        ++row;
    }
    if (row != REMC_NTYPES) {
        std::fprintf(stderr, "Pair table %s has %d rows, expected %d\n", filename, row, REMC_NTYPES);
        return false;
    }
    return true;
}

// --------------------------------------------------------- molecules ----

static std::string field(const std::string &line, size_t a, size_t b)
{
    if (line.size() <= a) return std::string();
    std::string s = line.substr(a, b - a);
    size_t i = s.find_first_not_of(' '), j = s.find_last_not_of(' ');
    return i == std::string::npos ? std::string() : s.substr(i, j - i + 1);
}

bool Molecule::initFromPDB(const char *pdbfilename, const AminoAcids &aas)
{
    // Molecule.cpp:430-830: Calpha ATOM records (name cols 13-16, resName
    // 18-20, chainID 22, resSeq 23-26, x/y/z 31-54); the centre is the float
    // mean of the positions and residues keep positions relative to it.
    std::vector<std::string> lines;
    if (!read_lines(pdbfilename, lines)) {
        std::fprintf(stderr, "Failed to open PDB file %s\n", pdbfilename);
        return false;
    }
    Residues.clear();
    for (const std::string &ln : lines) {
        if (ln.compare(0, 4, "ATOM") != 0 || (ln.size() > 4 && ln[4] != ' ')) continue;
        if (field(ln, 12, 16) != "CA") continue;
        std::string res = field(ln, 17, 20);
        int idx = aas.getAminoAcidIndex(res.c_str());
        if (idx < 0) {
            std::fprintf(stderr, "%s: unknown residue %s\n", pdbfilename, res.c_str());
            return false;
        }
        Residue r;
        r.aminoAcidIndex = idx;
        r.electrostaticCharge = aas.data[idx].electrostaticCharge;
        r.vanderWaalRadius = aas.data[idx].vanderWaalRadius;
        r.chainId = ln.size() > 21 ? ln[21] : ' ';
        r.resSeq = std::atoi(field(ln, 22, 26).c_str());
        r.position.x = (float)std::strtod(field(ln, 30, 38).c_str(), nullptr);
        r.position.y = (float)std::strtod(field(ln, 38, 46).c_str(), nullptr);
        r.position.z = (float)std::strtod(field(ln, 46, 54).c_str(), nullptr);
        Residues.push_back(r);
    }
    if (Residues.empty()) {
        std::fprintf(stderr, "%s: no CA atoms\n", pdbfilename);
        return false;
    }
    Vector3f acc;
    for (const Residue &r : Residues) {
        acc.x += r.position.x;
        acc.y += r.position.y;
        acc.z += r.position.z;
    }
    const float n = (float)Residues.size();
    center = Vector3f(acc.x / n, acc.y / n, acc.z / n);
    position = center;
    for (Residue &r : Residues)
        r.relativePosition = Vector3f(r.position.x - center.x, r.position.y - center.y, r.position.z - center.z);
    rotation = Quaternion();
    filename = pdbfilename;
    return true;
}

bool Molecule::translate(Vector3f v)
{
    // Molecule.cpp:146-171
    center.x += v.x;
    center.y += v.y;
    center.z += v.z;
    position.x += v.x;
    position.y += v.y;
    position.z += v.z;
    for (Residue &r : Residues) {
        r.position.x += v.x;
        r.position.y += v.y;
        r.position.z += v.z;
    }
    return true;
}

void Molecule::setPosition(Vector3f v)
{
    // Molecule.cpp:247-260
    center = v;
    position = v;
    for (Residue &r : Residues) {
        r.position.x = r.relativePosition.x + center.x;
        r.position.y = r.relativePosition.y + center.y;
        r.position.z = r.relativePosition.z + center.z;
    }
}

void Molecule::setRotation(Quaternion q)
{
    // Molecule.cpp:335-346: compose and rotate the relative positions
    q.normalize();
    rotation = q * rotation;
    for (Residue &r : Residues) {
        r.relativePosition = q.rotateVector(r.relativePosition);
        r.position.x = r.relativePosition.x + center.x;
        r.position.y = r.relativePosition.y + center.y;
        r.position.z = r.relativePosition.z + center.z;
    }
}

bool Molecule::rotateQ(Vector3double axis, double angle)
{
    // Molecule.cpp:288-333 (axis is a unit vector)
    const double s = std::sin(angle / 2.0), c = std::cos(angle / 2.0);
    setRotation(Quaternion(c, s * axis.x, s * axis.y, s * axis.z));
    return true;
}

void Molecule::saveBeforeStateChange(const Molecule *m)
{
    Residues = m->Residues;
    center = m->center;
    position = m->position;
    rotation = m->rotation;
}

void Molecule::undoStateChange(const Molecule *m)
{
    Residues = m->Residues;
    center = m->center;
    position = m->position;
    rotation = m->rotation;
}

void Molecule::setMoleculeRoleIdentifier(float role)
{
    // Molecule.cpp:118-139.  The reference also overwrites crowder residues'
    // types with -1; they are never read for crowder pairs, so types are kept.
    moleculeRoleIdentifier = role;
}

// ----------------------------------------------------------- replica ----

int Replica::loadMolecule(const char *pdbfilename)
{
    // Replica.cpp:230-254
    Molecule m;
    if (!m.initFromPDB(pdbfilename, aminoAcids)) return -1;
    m.index = (int)molecules.size();
    molecules.push_back(std::move(m));
    return (int)molecules.size() - 1;
}

int Replica::loadMolecule(const char *pdbfilename, Vector3f position, Vector3double rotationAxis,
                          double rotationAmount)
{
    // Replica.cpp:256-263
    int i = loadMolecule(pdbfilename);
    if (i < 0) return i;
    rotationAxis.normalizeInPlace();
    molecules[i].setPosition(position);
    molecules[i].rotateQ(rotationAxis, rotationAmount);
    return i;
}

int Replica::residueCount() const
{
    int n = 0;
    for (const Molecule &m : molecules) n += (int)m.Residues.size();
    return n;
}

long Replica::countpairs() const
{
    // inter-molecule residue pairs of one full evaluation
    long n = residueCount(), s = 0;
    for (const Molecule &m : molecules) s += (long)m.Residues.size() * (long)m.Residues.size();
    return (n * n - s) / 2;
}

void Replica::exchangeReplicas(Replica &r)
{
    // Replica.cpp:145-159: temperature, label and the statistics that follow
    // the temperature swap; conformations and step sizes stay.
    std::swap(label, r.label);
    std::swap(temperature, r.temperature);
    std::swap(totalAcceptReject, r.totalAcceptReject);
    std::swap(totalAccept, r.totalAccept);
    std::swap(boundSamples, r.boundSamples);
    std::swap(samplesSinceLastExchange, r.samplesSinceLastExchange);
    std::swap(totalBoundSamples, r.totalBoundSamples);
    std::swap(totalSamples, r.totalSamples);
}

remc_system_desc Replica::Packed::desc()
{
    remc_system_desc d;
    std::memset(&d, 0, sizeof d);
    d.n_beads = (int)x.size();
    d.n_bodies = (int)crowder.size();
    d.body_start = body_start.data();
    d.type = type.data();
    d.charge = charge.data();
    d.sigma = sigma.data();
    d.crowder = crowder.data();
    d.x = x.data();
    d.y = y.data();
    d.z = z.data();
    d.rx = rx.data();
    d.ry = ry.data();
    d.rz = rz.data();
    d.cx = cx.data();
    d.cy = cy.data();
    d.cz = cz.data();
    d.quat = quat.data();
    return d;
}

Replica::Packed Replica::pack() const
{
    Packed p;
    p.body_start.push_back(0);
    for (const Molecule &m : molecules) {
        for (const Residue &r : m.Residues) {
            p.type.push_back(r.aminoAcidIndex);
            p.charge.push_back(r.electrostaticCharge);
            p.sigma.push_back(r.vanderWaalRadius);
            p.x.push_back(r.position.x);
            p.y.push_back(r.position.y);
            p.z.push_back(r.position.z);
            p.rx.push_back(r.relativePosition.x);
            p.ry.push_back(r.relativePosition.y);
            p.rz.push_back(r.relativePosition.z);
        }
        p.body_start.push_back((int)p.x.size());
        p.crowder.push_back(m.isCrowder() ? 1 : 0);
        p.cx.push_back(m.center.x);
        p.cy.push_back(m.center.y);
        p.cz.push_back(m.center.z);
        p.quat.push_back(m.rotation.w);
        p.quat.push_back(m.rotation.x);
        p.quat.push_back(m.rotation.y);
        p.quat.push_back(m.rotation.z);
    }
    return p;
}

remc_forcefield makeForceField(const AminoAcids &aas)
{
    // definitions.h:11-35: constants evaluated in float as the reference's
    // float expressions are, widened to double.
    remc_forcefield ff;
    std::memset(&ff, 0, sizeof ff);
    for (int i = 0; i < REMC_NTYPES; ++i)
        for (int j = 0; j < REMC_NTYPES; ++j) ff.table[i * REMC_NTYPES + j] = aas.LJpotentials[i][j];
    volatile float lambda = 0.159f, e0 = -2.27f;
    volatile float kbt = 1.0f / (294.0f * 8.314472f / 4184.0f);
    volatile float e = 1.602176487f;
    volatile float dh = 0.948f * e * e;
    ff.lambda = lambda;
    ff.e0 = e0;
    ff.xi = 10.0;
    ff.lj_cf = 0.3507221006079f;
    ff.dh_cf = dh;
    ff.kbt_cf = kbt;
    ff.crowder_sigma = 6.0;
    ff.crowder_cutoff = 15.0;
    return ff;
}

// Owns one context; every failing call reports the library's message.
struct Ctx {
    remc_ctx *c = nullptr;
    explicit Ctx(int device)
    {
        if (remc_ctx_create(device, &c) != REMC_OK) c = nullptr;
    }
    ~Ctx()
    {
        if (c) remc_ctx_destroy(c);
    }
    Ctx(const Ctx &) = delete;
    Ctx &operator=(const Ctx &) = delete;
};

static bool ok(int rc, const char *what)
{
    if (rc == REMC_OK) return true;
    std::fprintf(stderr, "! %s failed (%d): %s\n", what, rc, remc_last_error());
    return false;
}

double Replica::E(int device)
{
    // Replica::EonDevice (Replica.cpp:1601-1616) on a context of its own
    Ctx ctx(device);
    if (!ctx.c) {
        ok(REMC_ERR_DEVICE, "remc_ctx_create");
        return NAN;
    }
    remc_forcefield ff = makeForceField(aminoAcids);
    Packed p = pack();
    remc_system_desc d = p.desc();
    d.box = boundingValue;
    int sys = -1, first = -1;
    double T = temperature;
    float ts = translateStep;
    double rs = rotateStep;
    double e = NAN;
    if (ok(remc_set_forcefield(ctx.c, &ff), "remc_set_forcefield") && ok(remc_system_add(ctx.c, &d, &sys), "remc_system_add") &&
        ok(remc_replicas_add(ctx.c, sys, 1, &T, &ts, &rs, &first), "remc_replicas_add") &&
        ok(remc_energy_all(ctx.c, &e), "remc_energy_all"))
        potential = e;
    return e;
}

int checkConformations(const AminoAcids &aas, const argdata &p)
{
    // main.cpp:1652-1707 sums each example by Replica::E() on the host and by
    // EonDevice; here the second device path, the batched full-energy kernel
    // (remc_energy_all), stands where the host sum stood.
    std::printf("performing check...\n");
    const float box = 118.4f;
    for (int i = 1; i <= 10; ++i) {
        Replica r;
        r.setAminoAcidData(aas);
        r.setLabel(i);
        r.setBoundingValue(box);
        char a[4096], b[4096];
        std::snprintf(a, sizeof a, "%s/conf%d/1a.pdb", p.checkDir.c_str(), i);
        std::snprintf(b, sizeof b, "%s/conf%d/1b.pdb", p.checkDir.c_str(), i);
        if (r.loadMolecule(a) < 0 || r.loadMolecule(b) < 0) {
            std::printf("! cannot load example conformation %d (%s, %s)\n", i, a, b);
            return 1;
        }
        r.nonCrowderCount = (int)r.molecules.size();
        const double batched = r.E(p.device);
        const double plain = r.EonDevice();
        r.FreeDevice();
        if (batched != batched || plain != plain) return 1;
        std::printf("PLAIN,BATCHED,err: %24.20f %24.20f %24.20f\n", plain, batched, std::fabs(plain - batched));
    }
    std::fflush(stdout);
    return 0;
}

double checkedInitialEnergy(Replica *initial, int device)
{
    const double e = initial->E(device);
    if (e != e) return e;
    const double plain = initial->EonDevice();
    initial->FreeDevice();
    if (plain != plain) return plain;
    std::printf("GPU initial energy value: %26.23f kcal/mol\n", e);
    std::printf("Absolute error:           %26.23f kcal/mol\n", std::fabs(e - plain));
    std::printf("Relative error:           %26.23f kcal/mol\n", std::fabs(plain - e) / std::fabs(plain));
    if (std::fabs(plain - e) / std::fabs(plain) > 0.01)
        std::printf("!\n! WARNING: Inconsistency between GPU and CPU result, check simulation \n!\n");
    return e;
}

void Replica::DeviceMirror::reset()
{
    if (ctx) remc_ctx_destroy(ctx);
    ctx = nullptr;
    sys = -1;
    seen = remc_counters{};
}

void Replica::copy(const Replica &r)
{
    label = r.label;
    temperature = r.temperature;
    aminoAcids = r.aminoAcids;
    boundingValue = r.boundingValue;
    nonCrowderCount = r.nonCrowderCount;
    molecules = r.molecules;
    potential = r.potential;
    blockSize = r.blockSize;
}

void Replica::initRNGs(uint64_t seed) { rngSeed = seed; }

int Replica::ReplicaDataToDevice(int device)
{
    mirror.reset();
    remc_ctx *c = nullptr;
    if (!ok(remc_ctx_create(device, &c), "remc_ctx_create")) return -1;
    mirror.ctx = c;
    remc_forcefield ff = makeForceField(aminoAcids);
    Packed p = pack();
    remc_system_desc d = p.desc();
    d.box = boundingValue;
    double T = temperature;
    float ts = translateStep;
    double rs = rotateStep;
    int first = -1;
    // one stream family per (seed, label): the label spreads replicas of one seed apart
    const uint64_t key = rngSeed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(label + 1));
    if (!ok(remc_set_forcefield(c, &ff), "remc_set_forcefield") || !ok(remc_set_seed(c, key), "remc_set_seed") ||
        !ok(remc_system_add(c, &d, &mirror.sys), "remc_system_add") ||
        !ok(remc_replicas_add(c, mirror.sys, 1, &T, &ts, &rs, &first), "remc_replicas_add")) {
        mirror.reset();
        return -1;
    }
    return 0;
}

int Replica::MoleculeDataToDevice(int moleculeId)
{
    if (moleculeId < 0 || moleculeId >= (int)molecules.size()) {
        std::fprintf(stderr, "MoleculeDataToDevice: molecule %d out of range\n", moleculeId);
        return -1;
    }
    if (!mirror.ctx) return ReplicaDataToDevice();
    // the whole conformation goes up (one copy per array); the tracked
    // energy is then re-evaluated
    Packed p = pack();
    if (!ok(remc_set_state(mirror.ctx, 0, p.x.data(), p.y.data(), p.z.data(), p.rx.data(), p.ry.data(), p.rz.data(),
                           p.cx.data(), p.cy.data(), p.cz.data(), p.quat.data()),
            "remc_set_state") ||
        !ok(remc_refresh_energies(mirror.ctx), "remc_refresh_energies"))
        return -1;
    return 0;
}

void Replica::FreeDevice() { mirror.reset(); }

double Replica::EonDevice()
{
    if (molecules.empty()) return 0.0;
    if (!mirror.ctx ? ReplicaDataToDevice() != 0 : MoleculeDataToDevice(0) != 0) return NAN;
    double e = NAN;
    if (!ok(remc_energy_plain(mirror.ctx, 0, blockSize, &e), "remc_energy_plain")) return NAN;
    return e;
}

int Replica::MCSearch(int steps)
{
    if (steps <= 0 || molecules.empty()) return 0;
    if (!mirror.ctx ? ReplicaDataToDevice() != 0 : MoleculeDataToDevice(0) != 0) return -1;
    remc_counters before{};
    if (!ok(remc_get_counters(mirror.ctx, &before), "remc_get_counters") ||
        !ok(remc_mc_steps(mirror.ctx, steps), "remc_mc_steps"))
        return -1;
    Packed p = pack();
    remc_counters after{};
    double e = 0.0;
    // potential is the full energy of the final conformation, as the reference's
    // MCSearch leaves it (the device tracks the sum of accepted dE, whose FP32
    // pair-term rounding accumulates over accepted steps)
    if (!ok(remc_get_state(mirror.ctx, 0, p.x.data(), p.y.data(), p.z.data(), p.rx.data(), p.ry.data(), p.rz.data(),
                           p.cx.data(), p.cy.data(), p.cz.data(), p.quat.data()),
            "remc_get_state") ||
        !ok(remc_refresh_energies(mirror.ctx), "remc_refresh_energies") ||
        !ok(remc_get_energies(mirror.ctx, &e), "remc_get_energies") ||
        !ok(remc_get_counters(mirror.ctx, &after), "remc_get_counters"))
        return -1;
    size_t k = 0;
    for (size_t m = 0; m < molecules.size(); ++m) {
        Molecule &mol = molecules[m];
        for (Residue &r : mol.Residues) {
            r.position = Vector3f(p.x[k], p.y[k], p.z[k]);
            r.relativePosition = Vector3f(p.rx[k], p.ry[k], p.rz[k]);
            ++k;
        }
        mol.center = Vector3f(p.cx[m], p.cy[m], p.cz[m]);
        mol.position = mol.center;
        mol.rotation = Quaternion(p.quat[4 * m], p.quat[4 * m + 1], p.quat[4 * m + 2], p.quat[4 * m + 3]);
    }
    potential = e;
    accept += (long)(after.accept - before.accept);
    acceptA += (long)(after.accept_boltz - before.accept_boltz);
    reject += (long)(after.reject - before.reject);
    return 0;
}

bool Replica::sample(FILE *boundConformations, int current_step, float boundEnergyThreshold, bool recordAllSamples)
{
    samplesSinceLastExchange++;
    float ncp = (float)potential;
    if ((int)molecules.size() != nonCrowderCount) {
        // the pairs between non-crowder molecules, on the device mirror
        if (!mirror.ctx ? ReplicaDataToDevice() != 0 : MoleculeDataToDevice(0) != 0) return false;
        double nc = 0.0;
        if (!ok(remc_energy_noncrowder_all(mirror.ctx, &nc), "remc_energy_noncrowder_all")) return false;
        ncp += (float)nc;
    }
    const bool bound = ncp < boundEnergyThreshold;
    if (bound) boundSamples++;
    if (boundConformations && (bound || recordAllSamples)) {
        std::fprintf(boundConformations, "%d; %0.5f (%0.5f); %0.1f\n", current_step, (double)ncp, potential,
                     (double)temperature);
        for (int a = 0; a < (int)molecules.size(); ++a) {
            const Molecule &m = molecules[a];
            std::fprintf(boundConformations, "%d %2d: %f %f %f %f %f %f %f\n", a < nonCrowderCount ? 0 : 1, a,
                         m.rotation.w, m.rotation.x, m.rotation.y, m.rotation.z, (double)m.center.x,
                         (double)m.center.y, (double)m.center.z);
        }
    }
    return bound;
}

void Replica::saveAsSinglePDB(const char *filename) const
{
    // Replica.cpp:495-530 record layout
    FILE *o = std::fopen(filename, "w");
    if (!o) {
        std::fprintf(stderr, "Cannot write %s\n", filename);
        return;
    }
    std::fprintf(o, "REMARK %s \n", filename);
    std::fprintf(o, "REMARK potential: %0.10f \n", (double)(float)potential);
    std::fprintf(o, "REMARK temperature: %5.1f \n", (double)temperature);
    for (size_t i = 0; i < molecules.size(); ++i) {
        const Molecule &m = molecules[i];
        std::fprintf(o, "REMARK Molecule: %d\n", (int)i);
        std::fprintf(o, "REMARK Rotation relative to input Q(w,x,y,z): %f %f %f %f\n", m.rotation.w, m.rotation.x,
                     m.rotation.y, m.rotation.z);
        std::fprintf(o, "REMARK Centriod position P(x,y,z): %f %f %f\n", (double)m.center.x, (double)m.center.y,
                     (double)m.center.z);
    }
    char chain = 'A';
    int item = 0;
    for (const Molecule &m : molecules) {
        int last = 0;
        for (const Residue &r : m.Residues) {
            ++item;
            std::fprintf(o, "ATOM  %5d %4s%c%3s %c%4d%c  %8.3f%8.3f%8.3f%6.2f%6.2f\n", item, "CA", ' ',
                         aminoAcids.data[r.aminoAcidIndex].sname.c_str(), chain, r.resSeq, ' ', (double)r.position.x,
                         (double)r.position.y, (double)r.position.z, 1.0, 1.0);
            last = r.resSeq;
        }
        if (!m.Residues.empty())
            std::fprintf(o, "TER   %5d      %3s %c%4d \n", item,
                         aminoAcids.data[m.Residues.back().aminoAcidIndex].sname.c_str(), chain, last);
        ++chain;
    }
    std::fprintf(o, "END \n");
    std::fclose(o);
}

// --------------------------------------------------- CLI and config ----

static void printHelp()
{
    std::printf("Usage: remc_sim -f <filename> [-h] [-q] [-t x] [-s x] [-g x] [-m x] [-e x] [-r x] [-o x] "
                "[-b x] [-t0 x] [-t1 x] [-bx x] [-c] [--seed x] [--device x] [--pid x] [--outdir d] [--data d]\n");
    std::printf("\t-h|--help: show this dialog\n");
    std::printf("\t-f|--file <file>: Input config file\n");
    std::printf("\t-q|--nosim:       Do everything except the simulation\n");
    std::printf("\t-c:              Check: the 10 example conformations summed by the plain and the batched kernel\n");
    std::printf("\t--check-dir d:   directory of the example conformations conf1..conf10 (default: data)\n");
    std::printf("The following values override those in the config file\n");
    std::printf("\t-t|--threads x:  accepted for compatibility (one device context runs every replica)\n");
    std::printf("\t-s|--streams x:  accepted for compatibility\n");
    std::printf("\t-g|--gpus x:     number of GPUs (1)\n");
    std::printf("\t-m|--mcsteps x:  How many MC Steps to perform per replica\n");
    std::printf("\t-e|--resteps x:  How many replica exhanges to perform\n");
    std::printf("\t-r|--replicas x: How many replicas\n");
    std::printf("\t-o|--output x:   log file name\n");
    std::printf("\t-b|--boundry x:  The bounding box edge length\n");
    std::printf("\t-t0|--tmin x:    The temperature of the lowest replica\n");
    std::printf("\t-t1|--tmax x:    The temperature of the highest replica\n");
    std::printf("\t-bx|--blockdim x: MC kernel workgroup size (256, 512 or 1024; default autotuned)\n");
    std::printf("\t--seed x:        seed of the counter-based RNG (default: time and pid)\n");
    std::printf("\t--device x:      HIP device index\n");
    std::printf("\t--pid x:         number used in output file names (default: process id)\n");
    std::printf("\t--outdir d:      directory of the output files (default: output)\n");
    std::printf("\t--data d:        directory holding AminoAcids and pair_potentials_MJ1996\n");
    std::printf("\t--save-pdb f:    write the initial conformation as a single PDB file\n");
    std::printf("\t--final-pdb f:   write the lowest-temperature replica's final conformation\n");
    std::printf("\t--coop x:        MC workgroups per replica (1, 2, 4 or 8; default autotuned)\n");
    std::printf("\t--checkpoint x:  write a checkpoint every x MC steps (rounded to exchange rounds)\n");
    std::printf("\t--checkpoint-dir d: directory of the checkpoint files (default: checkpoints)\n");
    std::printf("\t-z|--resume f:   continue the run saved in checkpoint file f\n");
}

bool getArgs(argdata *d, int argc, char **argv)
{
    // main.cpp:193-381.  Each option takes its value from the next argument.
    for (int i = 1; i < argc;) {
        const char *a = argv[i];
        auto is = [&](const char *s, const char *l) { return std::strcmp(a, s) == 0 || (l && std::strcmp(a, l) == 0); };
        auto val = [&]() -> const char * {
            if (i + 1 >= argc) {
                std::printf(" * Bad arguments\n");
                printHelp();
                return nullptr;
            }
            return argv[i + 1];
        };
        const char *v = nullptr;
        if (is("-h", "--help")) {
            printHelp();
            d->help = true;
            return true;
        } else if (is("-q", "--nosim")) {
            d->dryRun = true;
            ++i;
            continue;
        } else if (is("-v", "--view")) {
            std::printf("!!! This build does not support OpenGL and the -v option\n");
            ++i;
            continue;
        } else if (is("-c", nullptr)) {
            d->check = true;
            ++i;
            continue;
        }
        if (!(v = val())) return false;
        if (is("-t", "--threads")) d->threads = std::atoi(v);
        else if (is("-s", "--streams")) d->streams = std::atoi(v);
        else if (is("-g", "--gpus")) d->gpus = std::atoi(v);
        else if (is("-f", "--file")) {
            d->file = v;
            d->inputFile = true;
        } else if (is("-m", "--mcsteps")) d->MCsteps = std::atoi(v);
        else if (is("-e", "--resteps")) d->REsteps = std::atoi(v);
        else if (is("-r", "--replicas")) d->replicas = std::atoi(v);
        else if (is("-o", "--output")) d->logfile = v;
        else if (is("-b", "--boundry")) d->bound = (float)std::atof(v);
        else if (is("-t0", "--tmin")) d->temperatureMin = (float)std::atof(v);
        else if (is("-t1", "--tmax")) d->temperatureMax = (float)std::atof(v);
        else if (is("-bx", "--blockdim")) d->blockDim = std::atoi(v);
        else if (is("--seed", nullptr)) {
            d->seed = std::strtoull(v, nullptr, 0);
            d->seedGiven = true;
        } else if (is("--device", nullptr)) d->device = std::atoi(v);
        else if (is("--pid", nullptr)) d->pid = std::atoi(v);
        else if (is("--outdir", nullptr)) d->outputDir = v;
        else if (is("--data", nullptr)) d->dataDir = v;
        else if (is("--save-pdb", nullptr)) d->savePdb = v;
        else if (is("--check-dir", nullptr)) d->checkDir = v;
        else if (is("--final-pdb", nullptr)) d->finalPdb = v;
        else if (is("--coop", nullptr)) d->coop = std::atoi(v);
        else if (is("--checkpoint", nullptr)) d->checkpointFrequency = std::atoi(v);
        else if (is("--checkpoint-dir", nullptr)) d->checkpointDir = v;
        else if (is("-z", "--resume")) {
            d->resume = true;
            d->resumeFile = v;
        } else {
            std::printf(" * Bad arguments: %s\n", a);
            printHelp();
            return false;
        }
        i += 2;
    }
    return true;
}

static std::string resolve_path(const std::string &p, const std::string &config_dir)
{
    struct stat st;
    if (p.empty() || p[0] == '/' || stat(p.c_str(), &st) == 0 || config_dir.empty()) return p;
    std::string alt = config_dir + "/" + p;
    return stat(alt.c_str(), &st) == 0 ? alt : p;
}

static std::string out_path(const argdata *d, const char *what)
{
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s/%s_%d_%s", d->outputDir.c_str(), d->prependageString.c_str(), d->pid, what);
    return buf;
}

bool loadArgsFromFile(argdata *d, Replica *initial)
{
    // main.cpp:383-632: "key value" parameter lines, then a "files" section of
    // molecules and an optional "crowders" section, one per line as
    //   p(x,y,z) r(ax,ay,az,angle) file.pdb   (place the centre at x,y,z)
    //   t(x,y,z) r(ax,ay,az,angle) file.pdb   (translate by x,y,z)
    // with a rotation applied when angle >= 0.
    if (d->inputFile) {
        std::vector<std::string> lines;
        if (!read_lines(d->file.c_str(), lines)) {
            std::printf("Failed to open file: %s\n", d->file.c_str());
            return false;
        }
        std::string config_dir;
        {
            size_t s = d->file.find_last_of('/');
            if (s != std::string::npos) config_dir = d->file.substr(0, s);
        }
        bool params = true, crowders = false;
        int lc = 0;
        for (const std::string &raw : lines) {
            ++lc;
            std::string line = raw.substr(0, 511);
            char buf[512];
            std::snprintf(buf, sizeof buf, "%s", line.c_str());
            char *tok = std::strtok(buf, " ");
            if (!tok || line[0] == '#' || std::strlen(tok) == 0) continue;
            if (std::strcmp(tok, "files") == 0) {
                params = false;
                continue;
            }
            if (std::strcmp(tok, "crowders") == 0) {
                params = false;
                crowders = true;
                continue;
            }
            if (params) {
                char *v = std::strtok(nullptr, " ");
                if (!v) {
                    std::printf("Missing value on line %d: %s\n", lc, line.c_str());
                    continue;
                }
                if (!std::strcmp(tok, "gpus")) d->gpus = std::atoi(v);
                else if (!std::strcmp(tok, "streams")) d->streams = std::atoi(v);
                else if (!std::strcmp(tok, "threads")) d->threads = std::atoi(v);
                else if (!std::strcmp(tok, "mcsteps")) d->MCsteps = std::atoi(v);
                else if (!std::strcmp(tok, "resteps")) d->REsteps = std::atoi(v);
                else if (!std::strcmp(tok, "temperaturemax")) d->temperatureMax = (float)std::atof(v);
                else if (!std::strcmp(tok, "temperaturemin")) d->temperatureMin = (float)std::atof(v);
                else if (!std::strcmp(tok, "boundry")) d->bound = (float)std::atof(v);
                else if (!std::strcmp(tok, "replicas")) d->replicas = std::atoi(v);
                else if (!std::strcmp(tok, "samplefrequency")) d->sampleFrequency = std::atoi(v);
                else if (!std::strcmp(tok, "sampleafter")) d->sampleStartsAfter = std::atoi(v);
                else if (!std::strcmp(tok, "prefix")) d->prependageString = std::string(v) + "_repulsive";
                continue;
            }
            float px, py, pz, rx, ry, rz, ra;
            char fname[512];
            int got = 0;
            bool translate = true;
            if (line[0] == 't')
                got = std::sscanf(line.c_str(), "t(%f,%f,%f) r(%f,%f,%f,%f) %511s", &px, &py, &pz, &rx, &ry, &rz, &ra, fname);
            if (line[0] == 'p') {
                got = std::sscanf(line.c_str(), "p(%f,%f,%f) r(%f,%f,%f,%f) %511s", &px, &py, &pz, &rx, &ry, &rz, &ra, fname);
                translate = false;
            }
            if (got < 8) {
                std::printf("Failed to parse line %d: %s\n", lc, line.c_str());
                continue;
            }
            std::string path = resolve_path(fname, config_dir);
            int id = initial->loadMolecule(path.c_str());
            if (id < 0) return false;
            initial->molecules[id].filename = fname;
            if (translate) initial->molecules[id].translate(Vector3f(px, py, pz));
            else initial->molecules[id].setPosition(Vector3f(px, py, pz));
            if (ra >= 0.0f) {
                Vector3double ax(rx, ry, rz);
                ax.normalizeInPlace();
                initial->molecules[id].rotateQ(ax, ra);
            }
            if (crowders) initial->molecules[id].setMoleculeRoleIdentifier(CROWDER_IDENTIFIER);
            else d->nonCrowders++;
        }
    }
    if (d->bound <= 0) std::printf("! Bounding value too small, setting equal to %g\n", 150.0);
    initial->setBoundingValue(d->bound);
    if (d->temperatureMax < d->temperatureMin) {
        std::printf("! Maximum temperature < minimum temperature, swapping %g <-> %g\n", d->temperatureMax,
                    d->temperatureMin);
        std::swap(d->temperatureMax, d->temperatureMin);
    }
    if (d->threads > d->replicas) {
        std::printf("! Too many threads, setting equal to %d\n", d->replicas);
        d->threads = d->replicas;
    }
    initial->nonCrowderCount = d->nonCrowders;

    mkdir(d->outputDir.c_str(), 0755);
    std::string idx = out_path(d, "fileindex");
    FILE *fi = std::fopen(idx.c_str(), "w");
    if (!fi) {
        std::printf("Cannot create %s\n", idx.c_str());
        return false;
    }
    std::fprintf(fi, "index molecule_file_path crowder(Y/N)\n");
    std::printf("Argument data from file:\n-------------------------------------\n");
    std::printf("threads %d\nstreams %d\nGPUs %d\nmc steps %d\nre steps %d\nreplicas %d\n", d->threads, d->streams,
                d->gpus, d->MCsteps, d->REsteps, d->replicas);
    std::printf("sampling frequency (mc steps) %d\nsampling starts after (mc steps) %d\n", d->sampleFrequency,
                d->sampleStartsAfter);
    std::printf("bounding box size %g\nnon-crowder molecules %d\nmaximum temperature %g\nminimum temperature %g\n",
                d->bound, d->nonCrowders, d->temperatureMax, d->temperatureMin);
    std::printf("Loaded: \n-------------------------------------------------------------\n");
    for (int z = 0; z < (int)initial->molecules.size(); ++z) {
        const Molecule &m = initial->molecules[z];
        const bool crowder = z >= d->nonCrowders;
        std::printf("%2d %s %scentered @ (%0.3f,%0.3f,%0.3f)\n", z, m.filename.c_str(), crowder ? "crowder " : "",
                    (double)m.center.x, (double)m.center.y, (double)m.center.z);
        std::fprintf(fi, "%2d %s %c\n", z, m.filename.c_str(), crowder ? 'Y' : 'N');
    }
    std::printf("-------------------------------------------------------------\n");
    std::fclose(fi);
    return true;
}

// -------------------------------------------------------- checkpoint ----

bool saveSimulation(const std::string &path, const argdata &p, const Checkpoint &ck)
{
    // main.cpp:89-117 layout first, then the exact section
    const std::string tmp = path + ".tmp";
    FILE *o = std::fopen(tmp.c_str(), "w");
    if (!o) {
        std::printf("Cannot write checkpoint %s\n", tmp.c_str());
        return false;
    }
    const int R = (int)ck.energy.size(), N = ck.nBeads, M = ck.nBodies;
    std::fprintf(o, "#Saved simulation state file, prependage string, original pid\n");
    std::fprintf(o, "%s\n", p.file.c_str());
    std::fprintf(o, "%s_%d\n", p.prependageString.c_str(), p.pid);
    std::fprintf(o, "%d\n", p.pid);
    std::fprintf(o, "# threads, streams, gpus, MCsteps, RESteps, sampleFrequency, sampleStartsAfter,nonCrowders, "
                    "temperatureMin, temperatureMax, currentStep\n");
    std::fprintf(o, "%d %d %d %d %d %d %d %d %f %f %d\n", p.threads, p.streams, p.gpus, p.MCsteps, p.REsteps,
                 p.sampleFrequency, p.sampleStartsAfter, p.nonCrowders, (double)p.temperatureMin,
                 (double)p.temperatureMax, ck.roundsDone * (p.MCsteps / p.REsteps));
    std::fprintf(o, "#molecules\n%d\n#replicas\n%d\n", M, R);
    for (int r = 0; r < R; ++r) {
        std::fprintf(o, "#replica %d, molecules: x,y,z translation. w,x,y,z rotation quaternion\n", r);
        for (int m = 0; m < M; ++m) {
            const size_t b = (size_t)r * M + m;
            std::fprintf(o, "%.7f %.7f %.7f %.7f %.7f %.7f %.7f\n", (double)ck.cx[b], (double)ck.cy[b],
                         (double)ck.cz[b], ck.quat[4 * b], ck.quat[4 * b + 1], ck.quat[4 * b + 2], ck.quat[4 * b + 3]);
        }
        const SlotStats *st = nullptr;
        for (int k = 0; k < R; ++k)
            if (ck.ladder[k] == r) st = &ck.slots[k];
        std::fprintf(o, "#replica %d, stats: accept, acceptA, reject, totalAcceptReject, totalAccept, boundSamples, "
                        "samplesSinceLastExchange, totalBoundSamples, totalSamples\n", r);
        std::fprintf(o, "%ld %ld %ld %ld %ld %ld %ld %ld %ld\n", (long)(ck.counters[r].accept - ck.seen[r].accept),
                     (long)(ck.counters[r].accept_boltz - ck.seen[r].accept_boltz),
                     (long)(ck.counters[r].reject - ck.seen[r].reject), st->totalAcceptReject, st->totalAccept,
                     st->boundSamples, st->samplesSinceLastExchange, st->totalBoundSamples, st->totalSamples);
    }
    std::fprintf(o, "#exact: seed rounds_done total_exchanges total_tests replicas box beads bodies\n");
    std::fprintf(o, "%llu %d %d %d %d %a %d %d\n", (unsigned long long)ck.seed, ck.roundsDone, ck.totalExchanges,
                 ck.totalTests, R, (double)ck.box, N, M);
    std::fprintf(o, "#exact: ladder (replica on each rung), exchange round offset swaps\n");
    for (int k = 0; k < R; ++k) std::fprintf(o, "%d%c", ck.ladder[k], k + 1 < R ? ' ' : '\n');
    std::fprintf(o, "%llu %d %llu\n", (unsigned long long)ck.exRound, ck.exOffset, (unsigned long long)ck.exSwaps);
    std::fprintf(o, "#exact: per replica energy, accept accept_boltz reject rng_steps, the same at the last report\n");
    for (int r = 0; r < R; ++r)
        std::fprintf(o, "%a %llu %llu %llu %llu %llu %llu %llu\n", ck.energy[r],
                     (unsigned long long)ck.counters[r].accept, (unsigned long long)ck.counters[r].accept_boltz,
                     (unsigned long long)ck.counters[r].reject, (unsigned long long)ck.counters[r].steps,
                     (unsigned long long)ck.seen[r].accept, (unsigned long long)ck.seen[r].accept_boltz,
                     (unsigned long long)ck.seen[r].reject);
    std::fprintf(o, "#exact: per rung totalAccept totalAcceptReject boundSamples samplesSinceLastExchange "
                    "totalBoundSamples totalSamples\n");
    for (const SlotStats &st : ck.slots)
        std::fprintf(o, "%ld %ld %ld %ld %ld %ld\n", st.totalAccept, st.totalAcceptReject, st.boundSamples,
                     st.samplesSinceLastExchange, st.totalBoundSamples, st.totalSamples);
    for (int r = 0; r < R; ++r) {
        std::fprintf(o, "#exact replica %d: beads x y z rx ry rz, then bodies cx cy cz qw qx qy qz\n", r);
        for (int i = 0; i < N; ++i) {
            const size_t b = (size_t)r * N + i;
            std::fprintf(o, "%a %a %a %a %a %a\n", (double)ck.x[b], (double)ck.y[b], (double)ck.z[b], (double)ck.rx[b],
                         (double)ck.ry[b], (double)ck.rz[b]);
        }
        for (int m = 0; m < M; ++m) {
            const size_t b = (size_t)r * M + m;
            std::fprintf(o, "%a %a %a %a %a %a %a\n", (double)ck.cx[b], (double)ck.cy[b], (double)ck.cz[b],
                         ck.quat[4 * b], ck.quat[4 * b + 1], ck.quat[4 * b + 2], ck.quat[4 * b + 3]);
        }
    }
    const bool good = std::fflush(o) == 0 && !std::ferror(o);
    std::fclose(o);
    if (!good || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::printf("Cannot write checkpoint %s\n", path.c_str());
        return false;
    }
    return true;
}

namespace {

// Reads the non-comment lines of a checkpoint in order.
struct LineReader {
    std::vector<std::string> lines;
    size_t i = 0;
    const char *next()
    {
        while (i < lines.size() && (lines[i].empty() || lines[i][0] == '#')) ++i;
        return i < lines.size() ? lines[i++].c_str() : nullptr;
    }
};

bool parse_doubles(const char *s, double *out, int n)
{
    char *end = nullptr;
    for (int k = 0; k < n; ++k) {
        out[k] = std::strtod(s, &end);
        if (end == s) return false;
        s = end;
    }
    return true;
}

bool parse_u64s(const char *s, unsigned long long *out, int n)
{
    char *end = nullptr;
    for (int k = 0; k < n; ++k) {
        out[k] = std::strtoull(s, &end, 10);
        if (end == s) return false;
        s = end;
    }
    return true;
}

} // namespace

bool loadSimulation(const std::string &path, argdata *p, Checkpoint *ck)
{
    LineReader in;
    if (!read_lines(path.c_str(), in.lines)) {
        std::printf("Cannot open checkpoint %s\n", path.c_str());
        return false;
    }
    auto bad = [&](const char *what) {
        std::printf("Checkpoint %s: malformed %s\n", path.c_str(), what);
        return false;
    };
    const char *s;
    if (!(s = in.next())) return bad("config file line");
    p->file = s;
    p->inputFile = true;
    if (!(s = in.next())) return bad("prefix line");
    std::string prefix_pid = s;
    if (!(s = in.next())) return bad("pid line");
    p->pid = std::atoi(s);
    const std::string suffix = "_" + std::to_string(p->pid);
    if (prefix_pid.size() <= suffix.size() || prefix_pid.compare(prefix_pid.size() - suffix.size(), suffix.size(), suffix))
        return bad("prefix line");
    p->prependageString = prefix_pid.substr(0, prefix_pid.size() - suffix.size());
    if (!(s = in.next())) return bad("parameter line");
    int current = 0;
    if (std::sscanf(s, "%d %d %d %d %d %d %d %d %f %f %d", &p->threads, &p->streams, &p->gpus, &p->MCsteps, &p->REsteps,
                    &p->sampleFrequency, &p->sampleStartsAfter, &p->nonCrowders, &p->temperatureMin,
                    &p->temperatureMax, &current) != 11)
        return bad("parameter line");
    if (!(s = in.next())) return bad("molecule count");
    const int M = std::atoi(s);
    if (!(s = in.next())) return bad("replica count");
    const int R = std::atoi(s);
    if (M < 1 || R < 1) return bad("counts");
    // the reference-layout block (rounded poses, stats): the exact section below supersedes it
    for (int r = 0; r < R; ++r)
        for (int k = 0; k < M + 1; ++k)
            if (!in.next()) return bad("replica block");
    if (!(s = in.next())) return bad("exact header");
    unsigned long long seed = 0;
    int n_rep = 0, N = 0, Mx = 0;
    double box = 0;
    {
        char boxs[64];
        if (std::sscanf(s, "%llu %d %d %d %d %63s %d %d", &seed, &ck->roundsDone, &ck->totalExchanges, &ck->totalTests,
                        &n_rep, boxs, &N, &Mx) != 8)
            return bad("exact header");
        box = std::strtod(boxs, nullptr);
    }
    if (n_rep != R || Mx != M || N < 1) return bad("exact sizes");
    ck->seed = seed;
    ck->box = (float)box;
    ck->nBeads = N;
    ck->nBodies = M;
    p->seed = seed;
    p->seedGiven = true;
    p->replicas = R;
    p->bound = (float)box;
    if (!(s = in.next())) return bad("ladder");
    {
        ck->ladder.assign(R, 0);
        char *end = nullptr;
        for (int k = 0; k < R; ++k) {
            ck->ladder[k] = (int)std::strtol(s, &end, 10);
            if (end == s) return bad("ladder");
            s = end;
        }
    }
    if (!(s = in.next())) return bad("exchange state");
    {
        unsigned long long rd = 0, sw = 0;
        if (std::sscanf(s, "%llu %d %llu", &rd, &ck->exOffset, &sw) != 3) return bad("exchange state");
        ck->exRound = rd;
        ck->exSwaps = sw;
    }
    ck->energy.assign(R, 0.0);
    ck->counters.assign(R, remc_counters{});
    ck->seen.assign(R, remc_counters{});
    for (int r = 0; r < R; ++r) {
        if (!(s = in.next())) return bad("replica counters");
        char *end = nullptr;
        ck->energy[r] = std::strtod(s, &end);
        if (end == s) return bad("replica energy");
        unsigned long long v[7];
        if (!parse_u64s(end, v, 7)) return bad("replica counters");
        ck->counters[r] = remc_counters{v[0], v[1], v[2], v[3]};
        ck->seen[r] = remc_counters{v[4], v[5], v[6], 0};
    }
    ck->slots.assign(R, SlotStats{});
    for (int k = 0; k < R; ++k) {
        if (!(s = in.next())) return bad("rung statistics");
        SlotStats &st = ck->slots[k];
        if (std::sscanf(s, "%ld %ld %ld %ld %ld %ld", &st.totalAccept, &st.totalAcceptReject, &st.boundSamples,
                        &st.samplesSinceLastExchange, &st.totalBoundSamples, &st.totalSamples) != 6)
            return bad("rung statistics");
    }
    const size_t nb = (size_t)R * N, nm = (size_t)R * M;
    for (auto *v : {&ck->x, &ck->y, &ck->z, &ck->rx, &ck->ry, &ck->rz}) v->assign(nb, 0.f);
    for (auto *v : {&ck->cx, &ck->cy, &ck->cz}) v->assign(nm, 0.f);
    ck->quat.assign(4 * nm, 0.0);
    for (int r = 0; r < R; ++r) {
        for (int i = 0; i < N; ++i) {
            double v[6];
            if (!(s = in.next()) || !parse_doubles(s, v, 6)) return bad("bead line");
            const size_t b = (size_t)r * N + i;
            ck->x[b] = (float)v[0];
            ck->y[b] = (float)v[1];
            ck->z[b] = (float)v[2];
            ck->rx[b] = (float)v[3];
            ck->ry[b] = (float)v[4];
            ck->rz[b] = (float)v[5];
        }
        for (int m = 0; m < M; ++m) {
            double v[7];
            if (!(s = in.next()) || !parse_doubles(s, v, 7)) return bad("body line");
            const size_t b = (size_t)r * M + m;
            ck->cx[b] = (float)v[0];
            ck->cy[b] = (float)v[1];
            ck->cz[b] = (float)v[2];
            for (int t = 0; t < 4; ++t) ck->quat[4 * b + t] = v[3 + t];
        }
    }
    p->resume = true;
    return true;
}

// ------------------------------------------------------- REMC driver ----

namespace {

struct SamplingFiles {
    FILE *fractionBound = nullptr, *boundConformations = nullptr, *acceptanceRatio = nullptr, *exchangeFrequency = nullptr;
    ~SamplingFiles()
    {
        for (FILE *f : {fractionBound, boundConformations, acceptanceRatio, exchangeFrequency})
            if (f) std::fclose(f);
    }
};

FILE *open_append(const std::string &path, const char *mode)
{
    FILE *f = std::fopen(path.c_str(), mode);
    if (!f) f = std::fopen(path.c_str(), "wt");
    if (!f) std::printf("Cannot open/create file: %s\n", path.c_str());
    return f;
}

bool initSamplingFiles(const argdata *d, SamplingFiles &s, bool headers)
{
    // main.cpp:634-692: file names and header lines (a resumed run appends
    // to the files of the run it continues, without new headers)
    mkdir(d->outputDir.c_str(), 0755);
    if (!(s.fractionBound = open_append(out_path(d, "fractionBound"), "at"))) return false;
    if (headers) std::fprintf(s.fractionBound, "# Fraction Bound\n#Iteration InstantaniousAve CumulativeAve\n");
    if (!(s.boundConformations = open_append(out_path(d, "boundconformations"), "a+"))) return false;
    if (headers)
        std::fprintf(s.boundConformations,
                     "# iteration; complex free energy (replica free energy); temperature;\n# molecule: rotation(w,x,y,z) "
                     "position(x,y,z) 1 line per other molecule in the bound state\n");
    if (!(s.acceptanceRatio = open_append(out_path(d, "acceptance_ratios"), "a+"))) return false;
    if (headers) std::fprintf(s.acceptanceRatio, "# Iteration AcceptanceRatio\n");
    if (!(s.exchangeFrequency = open_append(out_path(d, "exchange_freq"), "a+"))) return false;
    if (headers) std::fprintf(s.exchangeFrequency, "# Iteration continous_ratio, inst_ratio\n");
    std::fflush(nullptr);
    return true;
}

} // namespace

int REMCSimulation(Replica *initial, argdata *p, std::vector<Replica> *final_replicas, const Checkpoint *resume)
{
    // main.cpp:1106-1567 on one device context: the replicas' MC rounds run
    // as batched launches, exchange rounds on the device; sampling, the
    // statistics and the output files follow the reference's schedule.
    if (initial->molecules.empty()) {
        std::printf("-!- No molecules loaded. -!-\n-!- Aborting REMCSimulation. -!- \n");
        return 1;
    }
    const int R = p->replicas;
    if (R < 1 || p->REsteps < 1 || p->MCsteps < 1 || p->sampleFrequency < 1) {
        std::printf("! replicas, resteps, mcsteps and samplefrequency must be positive\n");
        return 1;
    }
    if (p->gpus > 1) std::printf("! %d GPUs requested: this driver runs every replica on one device\n", p->gpus);

    // temperature and step-size ladders (main.cpp:1161-1188)
    std::vector<Replica> rep(R, *initial);
    const double gT = R > 1 ? std::pow((double)(p->temperatureMax / p->temperatureMin), 1.0 / (double)(R - 1)) : 1.0;
    const double gTr = R > 1 ? std::pow((double)(MAX_TRANSLATION / MIN_TRANSLATION), 1.0 / (double)(R - 1)) : 1.0;
    const double gRo = R > 1 ? std::pow((double)(MAX_ROTATION / MIN_ROTATION), 1.0 / (double)(R - 1)) : 1.0;
    std::vector<double> T(R), rs(R);
    std::vector<float> ts(R);
    for (int i = 0; i < R; ++i) {
        rep[i].setLabel(i);
        rep[i].setTemperature((float)(p->temperatureMin * std::pow(gT, i)));
        rep[i].setRotateStep(MIN_ROTATION * std::pow(gRo, i));
        rep[i].setTranslateStep((float)(MIN_TRANSLATION * std::pow(gTr, i)));
        std::printf("Replica %d %.3f %.3f %.3f\n", i, (double)rep[i].temperature, (double)rep[i].translateStep,
                    rep[i].rotateStep);
        T[i] = rep[i].temperature;
        ts[i] = rep[i].translateStep;
        rs[i] = rep[i].rotateStep;
    }
    for (int i = 1; i < R; ++i)
        if (!(T[i] > T[i - 1])) {
            std::printf("! the temperature ladder must be strictly ascending (tmin %g, tmax %g)\n", p->temperatureMin,
                        p->temperatureMax);
            return 1;
        }

    Ctx ctx(p->device);
    if (!ctx.c) {
        ok(REMC_ERR_DEVICE, "remc_ctx_create");
        return 1;
    }
    remc_forcefield ff = makeForceField(initial->aminoAcids);
    Replica::Packed packed = initial->pack();
    remc_system_desc desc = packed.desc();
    desc.box = p->bound;
    int sys = -1, first = -1;
    if (!ok(remc_set_forcefield(ctx.c, &ff), "remc_set_forcefield")) return 1;
    if (!ok(remc_set_seed(ctx.c, p->seed), "remc_set_seed")) return 1;
    if (!ok(remc_system_add(ctx.c, &desc, &sys), "remc_system_add")) return 1;
    if (!ok(remc_replicas_add(ctx.c, sys, R, T.data(), ts.data(), rs.data(), &first), "remc_replicas_add")) return 1;
    if (p->blockDim && !ok(remc_set_mc_block(ctx.c, p->blockDim), "remc_set_mc_block")) return 1;
    if (p->coop && !ok(remc_set_mc_coop(ctx.c, p->coop), "remc_set_mc_coop")) return 1;
    const int M = (int)initial->molecules.size(), N = (int)packed.x.size();

    // statistics that follow a temperature live with its ladder slot; the
    // accept/reject interval counters stay with the replica (Replica.cpp:145-159)
    std::vector<SlotStats> slot(R);
    std::vector<remc_counters> seen(R), now(R);
    std::memset(seen.data(), 0, sizeof(remc_counters) * R);
    std::vector<int> ladder(R);
    std::vector<double> energies(R), temps(R), noncrowder(R);
    int totalExchanges = 0, totalTests = 0, startRound = 1;

    if (resume) {
        // the saved run's state, replica by replica, and its bookkeeping
        const Checkpoint &ck = *resume;
        if ((int)ck.energy.size() != R || ck.nBeads != N || ck.nBodies != M || (int)ck.ladder.size() != R) {
            std::printf("! checkpoint sizes (%d replicas, %d beads, %d bodies) do not match the configuration\n",
                        (int)ck.energy.size(), ck.nBeads, ck.nBodies);
            return 1;
        }
        for (int r = 0; r < R; ++r) {
            const size_t b = (size_t)r * N, m = (size_t)r * M;
            if (!ok(remc_set_state(ctx.c, r, &ck.x[b], &ck.y[b], &ck.z[b], &ck.rx[b], &ck.ry[b], &ck.rz[b], &ck.cx[m],
                                   &ck.cy[m], &ck.cz[m], &ck.quat[4 * m]),
                    "remc_set_state"))
                return 1;
        }
        if (!ok(remc_set_energies(ctx.c, ck.energy.data()), "remc_set_energies")) return 1;
        if (!ok(remc_set_counters(ctx.c, ck.counters.data()), "remc_set_counters")) return 1;
        if (!ok(remc_set_ladder(ctx.c, sys, ck.ladder.data()), "remc_set_ladder")) return 1;
        if (!ok(remc_set_exchange_state(ctx.c, sys, ck.exRound, ck.exOffset, ck.exSwaps), "remc_set_exchange_state"))
            return 1;
        slot = ck.slots;
        seen = ck.seen;
        totalExchanges = ck.totalExchanges;
        totalTests = ck.totalTests;
        startRound = ck.roundsDone + 1;
        std::printf("--- Resuming after exchange round %d (MC step %d) ---\n", ck.roundsDone,
                    ck.roundsDone * (p->MCsteps / p->REsteps));
    }

    SamplingFiles files;
    if (!initSamplingFiles(p, files, !resume)) return 1;
    if (!resume) {
        std::fprintf(files.fractionBound, "Iteration:  ");
        std::fprintf(files.acceptanceRatio, "Iteration:  ");
        for (int k = 0; k < R; ++k) {
            std::fprintf(files.fractionBound, "%0.1fKi %0.1fKc ", T[k], T[k]);
            std::fprintf(files.acceptanceRatio, "%0.1fKi %0.1fKc ", T[k], T[k]);
        }
        std::fprintf(files.fractionBound, "\t\n");
        std::fprintf(files.acceptanceRatio, "\n");
    }

    int mcstepsPerRE = p->MCsteps / p->REsteps;
    if (mcstepsPerRE < 1) {
        std::printf("! mcsteps (%d) must be at least resteps (%d)\n", p->MCsteps, p->REsteps);
        return 1;
    }
    if (mcstepsPerRE % p->sampleFrequency > 0) {
        // main.cpp:1297-1311
        if (mcstepsPerRE > p->sampleFrequency) {
            while (mcstepsPerRE % p->sampleFrequency != 0) p->sampleFrequency++;
            std::printf("-!- CHANGED: sample frequency changed as it does divide into the replica exchange frequency. -!-\n");
            std::printf("-!- CHANGED: sample frequency <- %d. -!-\n", p->sampleFrequency);
        }
        if (mcstepsPerRE < p->sampleFrequency) {
            p->sampleFrequency = mcstepsPerRE;
            std::printf("-!- CHANGED: sample frequency too long for MC loop. -!-\n");
            std::printf("-!- CHANGED: sample frequency <- %d. -!-\n", mcstepsPerRE);
        }
    }
    const bool has_crowders = initial->nonCrowderCount != (int)initial->molecules.size();
    std::vector<float> cx(M), cy(M), cz(M);
    std::vector<double> quat(4 * M);

    auto slot_of = [&](std::vector<int> &slot_rep) {
        std::vector<int> s(R);
        for (int k = 0; k < R; ++k) s[slot_rep[k]] = k;
        return s;
    };

    // one sample of every replica (Replica::sample, Replica.cpp:1717-1780)
    auto sample_all = [&](int step) -> bool {
        if (!ok(remc_get_energies(ctx.c, energies.data()), "remc_get_energies")) return false;
        if (!ok(remc_get_temperatures(ctx.c, temps.data()), "remc_get_temperatures")) return false;
        if (!ok(remc_get_ladder(ctx.c, sys, ladder.data()), "remc_get_ladder")) return false;
        if (has_crowders && !ok(remc_energy_noncrowder_all(ctx.c, noncrowder.data()), "remc_energy_noncrowder_all"))
            return false;
        std::vector<int> s_of = slot_of(ladder);
        for (int r = 0; r < R; ++r) {
            SlotStats &st = slot[s_of[r]];
            st.samplesSinceLastExchange++;
            float ncp = (float)energies[r];
            if (has_crowders) ncp += (float)noncrowder[r];
            if (!(ncp < BOUND_ENERGY_VALUE)) continue;
            st.boundSamples++;
            if (!ok(remc_get_state(ctx.c, r, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, cx.data(), cy.data(),
                                   cz.data(), quat.data()),
                    "remc_get_state"))
                return false;
            std::fprintf(files.boundConformations, "%d; %0.5f (%0.5f); %0.1f\n", step, (double)ncp, energies[r],
                         (double)(float)temps[r]);
            for (int a = 0; a < M; ++a)
                std::fprintf(files.boundConformations, "%d %2d: %f %f %f %f %f %f %f\n",
                             a < initial->nonCrowderCount ? 0 : 1, a, quat[4 * a], quat[4 * a + 1], quat[4 * a + 2],
                             quat[4 * a + 3], (double)cx[a], (double)cy[a], (double)cz[a]);
        }
        return true;
    };

    // checkpoints fall on exchange rounds (main.cpp:1313-1325)
    int cpf = p->checkpointFrequency;
    if (cpf > 0 && cpf % mcstepsPerRE != 0)
        cpf = cpf < mcstepsPerRE ? mcstepsPerRE : (int)std::ceil(float(cpf) / float(mcstepsPerRE)) * mcstepsPerRE;
    const std::string cpath = p->checkpointDir + "/" + p->prependageString + "_" + std::to_string(p->pid);
    if (cpf > 0) {
        mkdir(p->checkpointDir.c_str(), 0755);
        std::printf("--- The simulation will checkpoint every %d monte carlo steps ---.\n", cpf);
    }
    auto checkpoint = [&](int roundsDone) -> bool {
        Checkpoint ck;
        ck.roundsDone = roundsDone;
        ck.totalExchanges = totalExchanges;
        ck.totalTests = totalTests;
        ck.seed = p->seed;
        ck.box = p->bound;
        ck.nBeads = N;
        ck.nBodies = M;
        ck.energy.assign(R, 0.0);
        ck.counters.assign(R, remc_counters{});
        ck.ladder.assign(R, 0);
        if (!ok(remc_get_energies(ctx.c, ck.energy.data()), "remc_get_energies") ||
            !ok(remc_get_counters(ctx.c, ck.counters.data()), "remc_get_counters") ||
            !ok(remc_get_ladder(ctx.c, sys, ck.ladder.data()), "remc_get_ladder") ||
            !ok(remc_get_exchange_state(ctx.c, sys, &ck.exRound, &ck.exOffset, &ck.exSwaps), "remc_get_exchange_state"))
            return false;
        ck.seen = seen;
        ck.slots = slot;
        const size_t nb = (size_t)R * N, nm = (size_t)R * M;
        for (auto *v : {&ck.x, &ck.y, &ck.z, &ck.rx, &ck.ry, &ck.rz}) v->assign(nb, 0.f);
        for (auto *v : {&ck.cx, &ck.cy, &ck.cz}) v->assign(nm, 0.f);
        ck.quat.assign(4 * nm, 0.0);
        for (int r = 0; r < R; ++r) {
            const size_t b = (size_t)r * N, m = (size_t)r * M;
            if (!ok(remc_get_state(ctx.c, r, &ck.x[b], &ck.y[b], &ck.z[b], &ck.rx[b], &ck.ry[b], &ck.rz[b], &ck.cx[m],
                                   &ck.cy[m], &ck.cz[m], &ck.quat[4 * m]),
                    "remc_get_state"))
                return false;
        }
        return saveSimulation(cpath, *p, ck);
    };

    std::printf("--- Staring simulation ---.\n");
    const int rounds = (p->MCsteps + mcstepsPerRE - 1) / mcstepsPerRE;
    int mcstep = (startRound - 1) * mcstepsPerRE;
    for (int round = startRound; round <= rounds; ++round) {
        for (int s = 0; s < mcstepsPerRE / p->sampleFrequency; ++s) {
            if (!ok(remc_mc_steps(ctx.c, p->sampleFrequency), "remc_mc_steps")) return 1;
            if (mcstep + s * p->sampleFrequency >= p->sampleStartsAfter && !sample_all(mcstep + s * p->sampleFrequency))
                return 1;
        }
        mcstep += mcstepsPerRE;
        if (round >= p->REsteps) continue;

        // exchange step `round` (main.cpp:1378-1469).  The tracked energies are
        // re-evaluated first: the reference's potential is a full evaluation of
        // the current conformation, and the running sum of accepted dE would
        // otherwise carry the FP32 rounding of every accepted step.
        if (!ok(remc_refresh_energies(ctx.c), "remc_refresh_energies")) return 1;
        if (!ok(remc_get_counters(ctx.c, now.data()), "remc_get_counters")) return 1;
        if (!ok(remc_get_ladder(ctx.c, sys, ladder.data()), "remc_get_ladder")) return 1;
        if (p->MCsteps / p->REsteps * (round - 1) >= p->sampleStartsAfter) {
            std::fprintf(files.fractionBound, "%9d: ", p->MCsteps / p->REsteps * round);
            std::fprintf(files.acceptanceRatio, "%9d: ", p->MCsteps / p->REsteps * round);
            for (int k = 0; k < R; ++k) {
                const int r = ladder[k];
                SlotStats &st = slot[k];
                float fb = float(st.boundSamples) / std::max(1.0f, float(st.samplesSinceLastExchange));
                st.totalBoundSamples += st.boundSamples;
                st.totalSamples += st.samplesSinceLastExchange;
                st.boundSamples = 0;
                st.samplesSinceLastExchange = 0;
                float cfb = float(st.totalBoundSamples) / float(st.totalSamples);
                std::fprintf(files.fractionBound, "| %6.4f %6.4f ", (double)fb, (double)cfb);
                const long acc = (long)(now[r].accept - seen[r].accept), accA = (long)(now[r].accept_boltz - seen[r].accept_boltz),
                           rej = (long)(now[r].reject - seen[r].reject);
                float ar = float(accA + acc) / float(accA + acc + rej);
                st.totalAccept += accA + acc;
                st.totalAcceptReject += accA + acc + rej;
                std::fprintf(files.acceptanceRatio, "| %6.4f %6.4f ", (double)ar,
                             (double)(float(st.totalAccept) / float(st.totalAcceptReject)));
                seen[r] = now[r];
            }
            std::fprintf(files.fractionBound, "\n");
            std::fprintf(files.acceptanceRatio, "\n");
            std::fflush(files.fractionBound);
            std::fflush(files.acceptanceRatio);
        }
        int swaps = 0;
        if (!ok(remc_exchange(ctx.c, &swaps), "remc_exchange")) return 1;
        // pairs (k, k+1) for k = offset, offset + 2, ... with the offset
        // alternating from 0: round r tests (R - 1 - (r-1)%2 + 1) / 2 pairs
        const int offset = (round - 1) % 2;
        const int tests = R - 1 > offset ? (R - 1 - offset + 1) / 2 : 0;
        totalExchanges += swaps;
        totalTests += tests;
        std::printf("Replica Exchange step %d of %d complete\n", round, p->REsteps);
        std::fprintf(files.exchangeFrequency, "%10d %6.4f %6.4f\n", round, (double)(float(totalExchanges) / float(totalTests)),
                     (double)(float(swaps) / float(tests)));
        std::fflush(files.exchangeFrequency);
        if (cpf > 0 && (round * mcstepsPerRE) % cpf == 0) {
            std::fflush(nullptr);
            if (!checkpoint(round)) return 1;
            std::printf("--- Checkpoint after exchange round %d: %s ---\n", round, cpath.c_str());
        }
    }
    std::printf("--- Replica Exchanges Complete.---\n--- All threads complete.---\n");

    // final lines (main.cpp:1489-1521), with the reference's final-line ratios
    if (!ok(remc_get_counters(ctx.c, now.data()), "remc_get_counters")) return 1;
    if (!ok(remc_get_ladder(ctx.c, sys, ladder.data()), "remc_get_ladder")) return 1;
    std::fprintf(files.fractionBound, "%9d: ", p->MCsteps);
    std::fprintf(files.acceptanceRatio, "%9d: ", p->MCsteps);
    for (int k = 0; k < R; ++k) {
        const int r = ladder[k];
        SlotStats &st = slot[k];
        float fb = float(st.boundSamples) / float(st.samplesSinceLastExchange);
        st.totalBoundSamples += st.boundSamples;
        st.totalSamples += st.samplesSinceLastExchange;
        st.boundSamples = 0;
        st.samplesSinceLastExchange = 0;
        float cfb = float(st.totalBoundSamples) / float(st.totalSamples);
        std::fprintf(files.fractionBound, "| %6.4f %6.4f  ", (double)fb, (double)cfb);
        const long acc = (long)(now[r].accept - seen[r].accept), accA = (long)(now[r].accept_boltz - seen[r].accept_boltz),
                   rej = (long)(now[r].reject - seen[r].reject);
        float ar = float(accA) / float(accA + acc + rej);
        std::fprintf(files.acceptanceRatio, "| %6.4f %6.4f ", (double)ar,
                     (double)(float(st.totalAccept) / float(st.totalAcceptReject)));
    }
    std::fprintf(files.fractionBound, "\n");
    std::fprintf(files.acceptanceRatio, "\n");
    std::printf("--- Simulation finished.---.\n\n");

    if (final_replicas) {
        // conformations, temperatures and statistics read back into Replica objects
        if (!ok(remc_get_energies(ctx.c, energies.data()), "remc_get_energies")) return 1;
        if (!ok(remc_get_temperatures(ctx.c, temps.data()), "remc_get_temperatures")) return 1;
        std::vector<float> x(N), y(N), z(N), rx(N), ry(N), rz(N);
        std::vector<int> s_of = slot_of(ladder);
        for (int r = 0; r < R; ++r) {
            if (!ok(remc_get_state(ctx.c, r, x.data(), y.data(), z.data(), rx.data(), ry.data(), rz.data(), cx.data(),
                                   cy.data(), cz.data(), quat.data()),
                    "remc_get_state"))
                return 1;
            Replica &o = rep[r];
            int b = 0;
            for (int a = 0; a < M; ++a) {
                Molecule &m = o.molecules[a];
                m.center = m.position = Vector3f(cx[a], cy[a], cz[a]);
                m.rotation = Quaternion(quat[4 * a], quat[4 * a + 1], quat[4 * a + 2], quat[4 * a + 3]);
                for (Residue &res : m.Residues) {
                    res.position = Vector3f(x[b], y[b], z[b]);
                    res.relativePosition = Vector3f(rx[b], ry[b], rz[b]);
                    ++b;
                }
            }
            o.potential = energies[r];
            o.temperature = (float)temps[r];
            o.label = s_of[r];
            o.accept = (long)(now[r].accept - seen[r].accept);
            o.acceptA = (long)(now[r].accept_boltz - seen[r].accept_boltz);
            o.reject = (long)(now[r].reject - seen[r].reject);
            const SlotStats &st = slot[s_of[r]];
            o.totalAccept = st.totalAccept;
            o.totalAcceptReject = st.totalAcceptReject;
            o.totalBoundSamples = st.totalBoundSamples;
            o.totalSamples = st.totalSamples;
        }
        *final_replicas = std::move(rep);
    }
    return 0;
}

} // namespace remc_facade

namespace remc_facade {

void Replica::writeToTrajectory(FILE *o, int frame, double drms) const
{
    // "post processor/postprocess/main.cpp:252-291" record layout
    std::fprintf(o, "TITLE     frame t=%7.3f\n", (double)(float)frame);
    std::fprintf(o, "MODEL%9d\n", frame);
    std::fprintf(o, "REMARK potential,drms: %0.10f %0.10f \n", potential, drms);
    char chain = 'A';
    int item = 0;
    for (const Molecule &m : molecules) {
        int last = 0;
        for (const Residue &r : m.Residues) {
            ++item;
            std::fprintf(o, "ATOM  %5d %4s%c%3s %c%4d%c  %8.3f%8.3f%8.3f%6.2f%6.2f\n", item, "CA", ' ',
                         aminoAcids.data[r.aminoAcidIndex].sname.c_str(), chain, r.resSeq, ' ', (double)r.position.x,
                         (double)r.position.y, (double)r.position.z, 1.0, 1.0);
            last = r.resSeq;
        }
        if (!m.Residues.empty())
            std::fprintf(o, "TER   %5d      %3s %c%4d \n", item,
                         aminoAcids.data[m.Residues.back().aminoAcidIndex].sname.c_str(), chain, last);
        ++chain;
    }
    std::fprintf(o, "ENDMDL\n");
    std::fflush(o);
}

double DRMS(const Replica &sim, const Replica &exp, bool reverseOrder)
{
    auto dist = [](const Vector3f &a, const Vector3f &b) {
        const double dx = (double)(a.x - b.x), dy = (double)(a.y - b.y), dz = (double)(a.z - b.z);
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    };
    const size_t M = sim.molecules.size();
    double drms = 0.0, n = 0.0;
    for (size_t a = 0; a < M; ++a)
        for (size_t b = a + 1; b < M; ++b)
            for (size_t i = 0; i < sim.molecules[a].Residues.size(); ++i)
                for (size_t j = 0; j < sim.molecules[b].Residues.size(); ++j) {
                    n += 1.0;
                    drms += std::fabs(dist(sim.molecules[a].Residues[i].position, sim.molecules[b].Residues[j].position) -
                                      dist(exp.molecules[a].Residues[i].position, exp.molecules[b].Residues[j].position));
                }
    drms /= n;
    bool same_sizes = true; // the swapped comparison pairs residue i of one molecule with residue i of the other
    for (size_t a = 0; a + 1 < M; ++a)
        same_sizes = same_sizes && sim.molecules[a].Residues.size() == sim.molecules[a + 1].Residues.size();
    if (reverseOrder && same_sizes) {
        double d2 = 0.0;
        for (size_t a = 0; a < M; ++a)
            for (size_t b = a + 1; b < M; ++b)
                for (size_t i = 0; i < sim.molecules[a].Residues.size(); ++i)
                    for (size_t j = 0; j < sim.molecules[b].Residues.size(); ++j)
                        d2 += std::fabs(
                            dist(sim.molecules[a].Residues[i].position, sim.molecules[b].Residues[j].position) -
                            dist(exp.molecules[b].Residues[i].position, exp.molecules[a].Residues[j].position));
        drms = std::min(drms, d2 / n);
    }
    return drms;
}

} // namespace remc_facade
