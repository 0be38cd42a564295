// Trajectory and energy-log writers.
//
// DcdWriter follows geodesic-io/src/dcd.rs (paths relative to the reference
// repository root): little-endian Fortran-record DCD, an 84-byte header block
// ("CORD", 20 control words: NSET (patched on close), ISTART 0, NSAVC =
// frame interval, NAMNF 0, DELTA = dt in AKMA units as f32, 24 in the last),
// one 80-character title line, NATOM, then per frame three records of f32
// x, y, z.  EnergyLogWriter follows export.rs:20-55: a CSV header, then
// step, time_ps (3 decimals), potential, kinetic, total (3 decimals),
// temperature 2 Ek / (3 N kB) in K (2 decimals), flushed per row.
#include <cstdio>
#include <cstring>

#include "geodesic.hpp"

namespace geodesic {
namespace {

constexpr double kAkmaPerPs = 20.455;           // dcd.rs:6
constexpr long kNsetOffset = 8;                 // dcd.rs:10
constexpr double kKbKcalPerMolK = 0.0019872041; // export.rs:9

FILE *as_file(void *p) { return static_cast<FILE *>(p); }

void put_i32(FILE *f, int32_t v) {
    unsigned char b[4];
    for (int k = 0; k < 4; k++) b[k] = (unsigned char)((uint32_t)v >> (8 * k));
    std::fwrite(b, 1, 4, f);
}

void put_f32(FILE *f, float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    put_i32(f, (int32_t)u);
}

}  // namespace

DcdWriter::DcdWriter(const std::string &path, size_t natom, uint32_t frame_interval, double dt_ps)
    : path_(path), natom_(natom) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw IoError(path, "cannot create file");
    file_ = f;
    const float delta = (float)(dt_ps * kAkmaPerPs);
    put_i32(f, 84);
    std::fwrite("CORD", 1, 4, f);
    for (int i = 0; i < 20; i++) {
        if (i == 2) put_i32(f, (int32_t)frame_interval);  // NSAVC
        else if (i == 9) put_f32(f, delta);               // DELTA
        else if (i == 19) put_i32(f, 24);                 // CHARMM version marker
        else put_i32(f, 0);                               // NSET (patched), ISTART, NAMNF, ...
    }
    put_i32(f, 84);
    char title[80];
    std::memset(title, ' ', sizeof(title));
    const char *msg = "Created by GEODESIC-M";
    std::memcpy(title, msg, std::strlen(msg));
    put_i32(f, 84);
    put_i32(f, 1);
    std::fwrite(title, 1, 80, f);
    put_i32(f, 84);
    put_i32(f, 4);
    put_i32(f, (int32_t)natom);
    put_i32(f, 4);
    if (std::ferror(f)) throw IoError(path, "write failed");
}

void DcdWriter::write_frame(const float *x, const float *y, const float *z, size_t n) {
    if (n != natom_)
        throw IoError(path_, "frame has " + std::to_string(n) + " values per axis, writer expects " +
                                 std::to_string(natom_) + " atoms");
    FILE *f = as_file(file_);
    for (const float *a : {x, y, z}) {
        put_i32(f, (int32_t)(n * 4));
        for (size_t i = 0; i < n; i++) put_f32(f, a[i]);
        put_i32(f, (int32_t)(n * 4));
    }
    if (std::ferror(f)) throw IoError(path_, "write failed");
    n_frames_++;
}

void DcdWriter::close() {
    if (closed_ || !file_) return;
    FILE *f = as_file(file_);
    closed_ = true;
    const bool ok = std::fseek(f, kNsetOffset, SEEK_SET) == 0;
    if (ok) put_i32(f, (int32_t)n_frames_);
    const bool bad = !ok || std::ferror(f) || std::fflush(f) != 0;
    std::fclose(f);
    file_ = nullptr;
    if (bad) throw IoError(path_, "cannot patch the frame count");
}

DcdWriter::~DcdWriter() {
    // best effort, as the reference's Drop: close() is the way to see errors
    try {
        close();
    } catch (...) {
    }
    if (file_) std::fclose(as_file(file_));
}

EnergyLogWriter::EnergyLogWriter(const std::string &path, size_t n_atoms) : path_(path), n_atoms_(n_atoms) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw IoError(path, "cannot create file");
    file_ = f;
    std::fputs("step,time_ps,potential_kcal,kinetic_kcal,total_kcal,temperature_K\n", f);
    if (std::fflush(f) != 0) throw IoError(path, "write failed");
}

void EnergyLogWriter::write_row(uint64_t step, double time_ps, double potential_kcal, double kinetic_kcal) {
    const double total = potential_kcal + kinetic_kcal;
    const double temperature = 2.0 * kinetic_kcal / (3.0 * (double)n_atoms_ * kKbKcalPerMolK);
    FILE *f = as_file(file_);
    std::fprintf(f, "%llu,%.3f,%.3f,%.3f,%.3f,%.2f\n", (unsigned long long)step, time_ps, potential_kcal,
                 kinetic_kcal, total, temperature);
    if (std::fflush(f) != 0) throw IoError(path_, "write failed");
}

EnergyLogWriter::~EnergyLogWriter() {
    if (file_) std::fclose(as_file(file_));
}

}  // namespace geodesic
