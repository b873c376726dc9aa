// The reference's text formats and its KMC event packet, on the host:
// DL_POLY-like basis / restart / archive (Basis::inputBasis, dumpBasis),
// the events.<n> file (writeEventList with Event::writeCentre/Saddle/
// Basin2) and the int[N] + float[6N+7] event packet (sendEvent,
// Event(int*, float*, ...)).  The writers use the same iostream manipulators
// as the reference, so the bytes match.
#include <math.h>

#include <algorithm>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "akmc.h"
#include "akmc_internal.h"

namespace {

std::vector<std::string> words_of(const std::string &s)
{
    std::istringstream is(s);
    std::vector<std::string> w;
    std::string t;
    while (is >> t) w.push_back(t);
    return w;
}

int io_fail(int code, const std::string &msg) { return akmc_fail(code, "%s", msg.c_str()); }

// the species names of a whitespace-separated list; label k is names[k]
bool name_of(const std::vector<std::string> &names, int label, std::string &out)
{
    if (label < 0 || label >= (int)names.size()) return false;
    out = names[label];
    return true;
}

void fixed10(std::ostream &o)
{
    o << std::setiosflags(std::ios::showpoint | std::ios::fixed) << std::setprecision(10);
}

// the three lattice lines of dumpBasis / printVectors (Common/Basis.cpp:623-640, 1091-1112)
void lattice_lines(std::ostream &o, const double *lat)
{
    for (int r = 0; r < 3; ++r) {
        fixed10(o);
        o << std::setw(15) << lat[3 * r] << std::setw(15) << lat[3 * r + 1] << std::setw(15) << lat[3 * r + 2]
          << std::endl;
    }
}

}  // namespace

extern "C" {

int akmc_io_dump_basis(const char *path, int append, int header, const char *title, const char *species, int n,
                       const int *label, const double *pos, const double lat[9], double simulation_time,
                       double energy, int iteration)
{
    if (!path || !species || n < 0 || (n > 0 && (!label || !pos)) || !lat) return io_fail(AKMC_EINVAL, "bad arguments");
    std::ofstream o(path, append ? std::ios::app : std::ios::trunc);
    if (!o) return io_fail(AKMC_EINVAL, std::string("could not open ") + path);
    const std::vector<std::string> names = words_of(species);
    // Basis::dumpBasis (Common/Basis.cpp:600-649); the header only on the
    // object's first write (firstWrite)
    if (header) {
        o << (title ? title : "") << std::endl;
        o << "0    3   " << n << std::endl;
    }
    fixed10(o);
    o << std::setw(15) << "timeStep " << iteration << "     " << n << "    0      3   " << simulation_time << "     "
      << energy << std::endl;
    lattice_lines(o, lat);
    for (int i = 0; i < n; ++i) {
        std::string s;
        if (!name_of(names, label[i], s)) return io_fail(AKMC_EINVAL, "atom label out of the species list");
        o << s << "      " << i + 1 << std::endl;
        fixed10(o);
        o << "  " << pos[i] << "  " << pos[n + i] << "  " << pos[2 * (size_t)n + i] << std::endl;
    }
    return o ? AKMC_SUCCESS : io_fail(AKMC_EINVAL, "write failed");
}

int akmc_io_write_restart(const char *path, int header, const char *title, const char *species, int n,
                          const int *label, const double *pos, const double lat[9], double simulation_time,
                          double energy, int iteration)
{
    if (!path) return io_fail(AKMC_EINVAL, "bad arguments");
    {
        // runKMC's restart file (kmc/KineticMonteCarlo.cpp:210-218)
        std::ofstream o(path, std::ios::trunc);
        if (!o) return io_fail(AKMC_EINVAL, std::string("could not open ") + path);
        o << "restart " << (title ? title : "") << std::endl;
        o << " 0   3 " << n << " " << n << std::endl;
    }
    return akmc_io_dump_basis(path, 1, header, title, species, n, label, pos, lat, simulation_time, energy,
                              iteration);
}

int akmc_io_basis_size(const char *path, int *n, int *levcfg)
{
    if (!path || !n) return io_fail(AKMC_EINVAL, "bad arguments");
    std::ifstream in(path);
    if (!in) return io_fail(AKMC_EINVAL, std::string("could not find basis file ") + path);
    std::string line;
    std::getline(in, line);
    if (!std::getline(in, line)) return io_fail(AKMC_EINVAL, "basis file has no size line");
    const auto w = words_of(line);
    try {
        if (levcfg) *levcfg = std::stoi(w.at(0));
        *n = std::stoi(w.at(2));
    } catch (const std::exception &) {
        return io_fail(AKMC_EINVAL, "malformed basis size line: " + line);
    }
    return AKMC_SUCCESS;
}

int akmc_io_read_basis(const char *path, const char *species, int n, int *label, double *pos, double lat[9],
                       int *iteration, double *time, char *title, int title_len)
{
    if (!path || !species || n < 0 || (n > 0 && (!label || !pos)) || !lat) return io_fail(AKMC_EINVAL, "bad arguments");
    std::ifstream in(path);
    if (!in) return io_fail(AKMC_EINVAL, std::string("could not find basis file ") + path);
    const std::vector<std::string> names = words_of(species);
    std::string line, t;
    // Basis::inputBasis (Common/Basis.cpp:356-501)
    std::getline(in, t);
    if (title && title_len > 0) {
        const size_t k = std::min(t.size(), (size_t)title_len - 1);
        t.copy(title, k);
        title[k] = 0;
    }
    try {
        std::getline(in, line);
        auto w = words_of(line);
        const int levcfg = std::stoi(w.at(0));
        if (std::stoi(w.at(2)) != n) return io_fail(AKMC_EINVAL, "atom count differs from the basis file's");
        if (iteration) *iteration = 0;
        if (time) *time = 0.0;
        std::getline(in, line);
        w = words_of(line);
        if (w.at(0) == "timeStep") {  // restart information, then the first lattice line
            if (iteration) *iteration = std::stoi(w.at(1));
            if (time) *time = std::stod(w.at(5));
            std::getline(in, line);
            w = words_of(line);
        }
        for (int r = 0; r < 3; ++r) {
            if (r > 0) {
                std::getline(in, line);
                w = words_of(line);
            }
            for (int c = 0; c < 3; ++c) lat[3 * r + c] = std::stod(w.at(c));
        }
        for (int j = 0; j < n; ++j) {
            if (!std::getline(in, line)) return io_fail(AKMC_EINVAL, "basis file ends early");
            w = words_of(line);
            int found = -1;
            for (int k = 0; k < (int)names.size(); ++k)
                if (w.at(0) == names[k]) found = k;
            if (found < 0) return io_fail(AKMC_EINVAL, "no atom of type " + w.at(0) + " exists");
            label[j] = found;
            std::getline(in, line);
            w = words_of(line);
            pos[j] = std::stod(w.at(0));
            pos[n + j] = std::stod(w.at(1));
            pos[2 * (size_t)n + j] = std::stod(w.at(2));
            if (levcfg >= 1) std::getline(in, line);
            if (levcfg >= 2) std::getline(in, line);
        }
    } catch (const std::exception &) {
        return io_fail(AKMC_EINVAL, "malformed basis file line: " + line);
    }
    return AKMC_SUCCESS;
}

int akmc_io_write_events(const char *path, const char *species, int n, const int *label, const double *basin1,
                         const double lat[9], double basin1_energy, int nev, const double *act_energy,
                         const double *centres, const double *saddle_energy, const double *saddles,
                         const double *basin2_energy, const double *basin2s, int chosen, double simulation_time,
                         int fp32_events)
{
    if (!path || !species || n < 0 || nev < 0 || !lat || (n > 0 && (!label || !basin1)) ||
        (nev > 0 && (!act_energy || !centres || !saddle_energy || !saddles || !basin2_energy || !basin2s)))
        return io_fail(AKMC_EINVAL, "bad arguments");
    const std::vector<std::string> names = words_of(species);
    // an event's quantities as the reference holds them: FP32 from the event
    // packet (kmc/Event.cpp:66-105) when fp32_events is set
    auto ev = [fp32_events](double v) { return fp32_events ? (double)(float)v : v; };
    std::ofstream o(path, std::ios::trunc);
    if (!o) return io_fail(AKMC_EINVAL, std::string("could not open ") + path);
    // writeEventList (kmc/KineticMonteCarlo.cpp:1561-1616)
    o << "events   " << nev << " chosen event " << chosen << " time " << simulation_time << std::endl;
    fixed10(o);
    o << std::setw(15) << "basin1 " << basin1_energy << std::endl;
    lattice_lines(o, lat);
    for (int i = 0; i < n; ++i) {
        std::string s;
        if (!name_of(names, label[i], s)) return io_fail(AKMC_EINVAL, "atom label out of the species list");
        fixed10(o);
        o << std::setw(5) << s << "    " << i << "\n"
          << std::setw(15) << basin1[i] << std::setw(15) << basin1[n + i] << std::setw(15) << basin1[2 * (size_t)n + i]
          << std::endl;
    }
    for (int e = 0; e < nev; ++e) {
        o << " event " << e << std::endl;
        o << " activation energy " << ev(act_energy[e]) << std::endl;
        // Event::writeCentre, writeSaddle, writeBasin2 (kmc/Event.cpp:220-290)
        o << std::setiosflags(std::ios::showpoint | std::ios::fixed) << std::setprecision(7) << std::setw(15)
          << "centre " << ev(centres[3 * e]) << " " << ev(centres[3 * e + 1]) << " " << ev(centres[3 * e + 2])
          << std::endl;
        for (int part = 0; part < 2; ++part) {
            const double *P = (part == 0 ? saddles : basin2s) + (size_t)e * 3 * n;
            const double E = part == 0 ? saddle_energy[e] : basin2_energy[e];
            fixed10(o);
            o << std::setw(15) << (part == 0 ? "saddle " : "basin2 ") << ev(E) << "  " << n << std::endl;
            for (int i = 0; i < n; ++i) {
                std::string s;
                name_of(names, label[i], s);
                fixed10(o);
                o << std::setw(5) << s << "    " << i << "\n"
                  << std::setw(15) << ev(P[i]) << std::setw(15) << ev(P[n + i]) << std::setw(15)
                  << ev(P[2 * (size_t)n + i]) << std::endl;
            }
        }
    }
    return o ? AKMC_SUCCESS : io_fail(AKMC_EINVAL, "write failed");
}

int akmc_kmc_event_pack(int n, const int *global_no, const double *basin2, const double *saddle, double e1,
                        double e2, double es, double ea, const double centre[3], int *veci, float *vecf)
{
    if (n < 0 || (n > 0 && (!global_no || !basin2 || !saddle || !veci)) || !centre || !vecf)
        return io_fail(AKMC_EINVAL, "bad arguments");
    // sendEvent (kmc/KineticMonteCarlo.cpp:1376-1461)
    for (int i = 0; i < n; ++i) {
        veci[i] = global_no[i];
        vecf[6 * i + 0] = (float)basin2[i];
        vecf[6 * i + 1] = (float)basin2[n + i];
        vecf[6 * i + 2] = (float)basin2[2 * (size_t)n + i];
        vecf[6 * i + 3] = (float)saddle[i];
        vecf[6 * i + 4] = (float)saddle[n + i];
        vecf[6 * i + 5] = (float)saddle[2 * (size_t)n + i];
    }
    const size_t d = 6 * (size_t)n;
    vecf[d] = (float)e1;
    vecf[d + 1] = (float)e2;
    vecf[d + 2] = (float)es;
    vecf[d + 3] = (float)ea;
    vecf[d + 4] = (float)centre[0];
    vecf[d + 5] = (float)centre[1];
    vecf[d + 6] = (float)centre[2];
    return AKMC_SUCCESS;
}

int akmc_kmc_event_unpack(int n, const int *veci, const float *vecf, double prefactor, double temperature,
                          int *global_no, double *basin2, double *saddle, double energies[4], double centre[3],
                          double *rate)
{
    if (n < 0 || (n > 0 && (!veci || !global_no || !basin2 || !saddle)) || !vecf || !energies || !centre)
        return io_fail(AKMC_EINVAL, "bad arguments");
    // Event::Event(int*, float*, natoms, prefactor, temperature) (kmc/Event.cpp:66-105)
    for (int i = 0; i < n; ++i) {
        basin2[i] = vecf[6 * i + 0];
        basin2[n + i] = vecf[6 * i + 1];
        basin2[2 * (size_t)n + i] = vecf[6 * i + 2];
        saddle[i] = vecf[6 * i + 3];
        saddle[n + i] = vecf[6 * i + 4];
        saddle[2 * (size_t)n + i] = vecf[6 * i + 5];
        global_no[i] = veci[i];
    }
    const size_t d = 6 * (size_t)n;
    for (int k = 0; k < 4; ++k) energies[k] = (double)vecf[d + k];
    for (int k = 0; k < 3; ++k) centre[k] = (double)vecf[d + 4 + k];
    if (rate) *rate = akmc_kmc_rate(energies[3], prefactor, temperature);
    return AKMC_SUCCESS;
}

}  // extern "C"
