// The reference's binary MAT file format (Mat-5.2/_extra/__File.h:1-66):
// a file holds int32 rows, int32 cols (the 8 bytes at &A.rows, :28 / :62),
// then rows * cols elements of the matrix's own type in its storage order,
// with no type tag and no padding.  The headerless forms (:13-19, :36-44,
// :49-55) hold the elements only.  Host-only code: these functions move
// bytes between a file and host memory and never touch a device.
//
// Differences from the reference, all on the error paths it leaves
// undefined: a file that cannot be opened, is shorter than its header says,
// or whose header disagrees with the requested shape is reported
// (MQR_ERR_IO) instead of leaving the destination half-filled.
#include <cstdio>
#include <cstring>

#include "mqr_internal.h"

namespace {

struct File {
    std::FILE *f = nullptr;
    File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

int64_t file_bytes(std::FILE *f) {
    if (std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long v = std::ftell(f);
    if (std::fseek(f, 0, SEEK_SET) != 0) return -1;
    return (int64_t)v;
}

bool elem_ok(size_t es) { return es == 1 || es == 2 || es == 4 || es == 8 || es == 16; }

}  // namespace

extern "C" {

int mqr_file_size(const char *path, int64_t *bytes) {
    if (!path) return -1;
    if (!bytes) return -2;
    *bytes = -1;
    File f(path, "rb");
    if (!f.f) return MQR_ERR_IO;
    *bytes = file_bytes(f.f);
    return *bytes < 0 ? MQR_ERR_IO : MQR_SUCCESS;
}

int mqr_mat_save(const char *path, int32_t rows, int32_t cols, const void *data, size_t elem_size) {
    if (!path) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const size_t count = (size_t)rows * (size_t)cols;
    if (count > 0 && !data) return -4;
    if (!elem_ok(elem_size)) return -5;
    File f(path, "wb");
    if (!f.f) return MQR_ERR_IO;
    const int32_t hdr[2] = {rows, cols};
    if (std::fwrite(hdr, sizeof(int32_t), 2, f.f) != 2) return MQR_ERR_IO;
    if (count > 0 && std::fwrite(data, elem_size, count, f.f) != count) return MQR_ERR_IO;
    if (std::fflush(f.f) != 0) return MQR_ERR_IO;
    return MQR_SUCCESS;
}

int mqr_mat_load_header(const char *path, int32_t *rows, int32_t *cols) {
    if (!path) return -1;
    if (!rows) return -2;
    if (!cols) return -3;
    File f(path, "rb");
    if (!f.f) return MQR_ERR_IO;
    int32_t hdr[2];
    if (std::fread(hdr, sizeof(int32_t), 2, f.f) != 2) return MQR_ERR_IO;
    if (hdr[0] < 0 || hdr[1] < 0) return MQR_ERR_IO;
    *rows = hdr[0];
    *cols = hdr[1];
    return MQR_SUCCESS;
}

int mqr_mat_load(const char *path, int32_t rows, int32_t cols, void *data, size_t elem_size) {
    if (!path) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const size_t count = (size_t)rows * (size_t)cols;
    if (count > 0 && !data) return -4;
    if (!elem_ok(elem_size)) return -5;
    File f(path, "rb");
    if (!f.f) return MQR_ERR_IO;
    const int64_t total = file_bytes(f.f);
    int32_t hdr[2];
    if (total < 8 || std::fread(hdr, sizeof(int32_t), 2, f.f) != 2) return MQR_ERR_IO;
    if (hdr[0] != rows || hdr[1] != cols) return MQR_ERR_IO;
    if ((uint64_t)total != 8 + (uint64_t)count * elem_size) return MQR_ERR_IO;
    if (count > 0 && std::fread(data, elem_size, count, f.f) != count) return MQR_ERR_IO;
    return MQR_SUCCESS;
}

int mqr_raw_save(const char *path, const void *data, size_t bytes) {
    if (!path) return -1;
    if (bytes > 0 && !data) return -2;
    File f(path, "wb");
    if (!f.f) return MQR_ERR_IO;
    if (bytes > 0 && std::fwrite(data, 1, bytes, f.f) != bytes) return MQR_ERR_IO;
    if (std::fflush(f.f) != 0) return MQR_ERR_IO;
    return MQR_SUCCESS;
}

int mqr_raw_load(const char *path, void *data, size_t bytes) {
    if (!path) return -1;
    if (bytes > 0 && !data) return -2;
    File f(path, "rb");
    if (!f.f) return MQR_ERR_IO;
    const int64_t total = file_bytes(f.f);
    if (total < 0 || (uint64_t)total < bytes) return MQR_ERR_IO;
    if (bytes > 0 && std::fread(data, 1, bytes, f.f) != bytes) return MQR_ERR_IO;
    return MQR_SUCCESS;
}

}  // extern "C"
