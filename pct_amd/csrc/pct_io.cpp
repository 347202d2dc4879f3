/*
 * pct_io.cpp -- PCTD version-0 projection files and text dumps (host code,
 * include/pct_io.h).
 *
 * Reader: the version-0 branches of count_histories_v0 (.cu:997-1097) and
 * read_data_chunk_v02 (.cu:1769-1933): header fields, then the 13 event
 * arrays in file order (t0..t3, v0..v3, u0..u3, WEPL) into the raw tracker
 * arrays (t/v/u in_1, in_2, out_1, out_2), mm -> cm as `value *= MM_TO_CM`
 * on a float with the double constant 0.1 (data.h:452, .cu:1863-1878), and
 * the gantry angle (int(projection_angle) + 270) % 360 for every event
 * (.cu:1917). Unlike the reference, a malformed or short file is an error
 * rather than a partial read.
 * Writer: the same layout, for synthetic scans.
 * Text dumps: array_2_disk (.cu:8092-8134), same file names, loop bounds and
 * stream formatting.
 * MLP path and path-error files of the CPU research loop (write_MLP_path,
 * read_MLP_path, read_MLP_path_error, .cu:4746-4761, 7479-7525).
 */
#include <stdio.h>
#include <string.h>
#include <fstream>
#include <string>

#include "../../include/pct_io.h"

namespace {

thread_local std::string g_err;

int io_fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

struct file_closer {
    FILE* f;
    ~file_closer() {
        if (f) fclose(f);
    }
};

bool read_exact(FILE* f, void* dst, size_t bytes) { return bytes == 0 || fread(dst, 1, bytes, f) == bytes; }

bool read_string(FILE* f, char* dst) {
    int32_t len = 0;
    if (!read_exact(f, &len, sizeof len) || len < 0 || len > (1 << 24)) return false;
    std::string s((size_t)len, '\0');
    if (!read_exact(f, &s[0], (size_t)len)) return false;
    size_t keep = s.size() < PCT_PCTD_STR - 1 ? s.size() : PCT_PCTD_STR - 1;
    memcpy(dst, s.data(), keep);
    dst[keep] = '\0';
    return true;
}

int read_header(FILE* f, const char* path, pct_pctd_header* h) {
    memset(h, 0, sizeof *h);
    char magic[4];
    if (!read_exact(f, magic, 4) || memcmp(magic, "PCTD", 4) != 0)
        return io_fail(PCT_E_IO, std::string(path) + ": not a PCTD file");
    if (!read_exact(f, &h->version, sizeof h->version)) return io_fail(PCT_E_IO, std::string(path) + ": short header");
    if (h->version != 0)
        return io_fail(PCT_E_IO, std::string(path) + ": PCTD version " + std::to_string(h->version) + " is not read");
    if (!read_exact(f, &h->n_histories, 4) || !read_exact(f, &h->projection_angle, 4) ||
        !read_exact(f, &h->beam_energy, 4) || !read_exact(f, &h->generation_date, 4) ||
        !read_exact(f, &h->preprocess_date, 4) || !read_string(f, h->phantom_name) ||
        !read_string(f, h->data_source) || !read_string(f, h->prepared_by))
        return io_fail(PCT_E_IO, std::string(path) + ": short header");
    if (h->n_histories < 0) return io_fail(PCT_E_IO, std::string(path) + ": negative event count");
    long off = ftell(f);
    if (off < 0) return io_fail(PCT_E_IO, std::string(path) + ": cannot tell the data offset");
    h->data_offset = (int64_t)off;
    return PCT_OK;
}

/* a byte mask printed as the reference's bool arrays print (0 / 1) */
struct mask_byte {
    uint8_t v;
};
std::ostream& operator<<(std::ostream& o, mask_byte m) { return o << (m.v != 0); }

template <class T>
void dump_values(std::ofstream& out, const T* data, int x_max, int y_max, int z_start, int z_end, int64_t elements) {
    int64_t index = 0;
    for (int z = z_start; z < z_end; z++) {
        for (int y = 0; y < y_max; y++) {
            for (int x = 0; x < x_max; x++) {
                index = x + ((int64_t)y * x_max) + ((int64_t)z * x_max * y_max);
                if (index >= elements) break;
                out << data[index] << " ";
            }
            if (index >= elements) break;
            out << std::endl;
        }
        if (index >= elements) break;
    }
}

}  // namespace

extern "C" const char* pct_io_last_error(void) { return g_err.c_str(); }

extern "C" int pct_pctd_read_header(const char* path, pct_pctd_header* header) {
    if (!path || !header) return io_fail(PCT_E_INVALID, "pct_pctd_read_header: null argument");
    file_closer fc{fopen(path, "rb")};
    if (!fc.f) return io_fail(PCT_E_IO, std::string(path) + ": cannot open");
    return read_header(fc.f, path, header);
}

extern "C" int pct_pctd_read(const char* path, int data_in_mm, int64_t capacity, const pct_synth_soa* out,
                             int64_t* n_read) {
    if (!path || !out || !n_read) return io_fail(PCT_E_INVALID, "pct_pctd_read: null argument");
    file_closer fc{fopen(path, "rb")};
    if (!fc.f) return io_fail(PCT_E_IO, std::string(path) + ": cannot open");
    pct_pctd_header h;
    int rc = read_header(fc.f, path, &h);
    if (rc) return rc;
    const int64_t n = h.n_histories;
    if (n > capacity)
        return io_fail(PCT_E_INVALID, std::string(path) + ": " + std::to_string(n) + " events exceed the capacity " +
                                          std::to_string(capacity));
    float* arrays[13] = {out->t_in_1, out->t_in_2, out->t_out_1, out->t_out_2, out->v_in_1, out->v_in_2, out->v_out_1,
                         out->v_out_2, out->u_in_1, out->u_in_2, out->u_out_1, out->u_out_2, out->wepl};
    if (n > 0) {
        for (float* a : arrays)
            if (!a) return io_fail(PCT_E_INVALID, "pct_pctd_read: null output array");
        if (!out->gantry_angle) return io_fail(PCT_E_INVALID, "pct_pctd_read: null gantry_angle array");
    }
    for (float* a : arrays)
        if (!read_exact(fc.f, a, sizeof(float) * (size_t)n))
            return io_fail(PCT_E_IO, std::string(path) + ": short event data");
    if (data_in_mm) {
        const double MM_TO_CM = 0.1;
        for (float* a : arrays)
            for (int64_t i = 0; i < n; ++i) a[i] *= MM_TO_CM;
    }
    const int32_t gantry = ((int)h.projection_angle + 270) % 360;
    for (int64_t i = 0; i < n; ++i) out->gantry_angle[i] = gantry;
    *n_read = n;
    return PCT_OK;
}

extern "C" int pct_pctd_write(const char* path, const pct_pctd_header* header, const pct_raw_soa* in) {
    if (!path || !header || !in) return io_fail(PCT_E_INVALID, "pct_pctd_write: null argument");
    if (in->n < 0 || in->n > INT32_MAX) return io_fail(PCT_E_INVALID, "pct_pctd_write: event count outside int32");
    const float* arrays[13] = {in->t_in_1, in->t_in_2, in->t_out_1, in->t_out_2, in->v_in_1, in->v_in_2, in->v_out_1,
                               in->v_out_2, in->u_in_1, in->u_in_2, in->u_out_1, in->u_out_2, in->wepl};
    if (in->n > 0)
        for (const float* a : arrays)
            if (!a) return io_fail(PCT_E_INVALID, "pct_pctd_write: null input array");
    file_closer fc{fopen(path, "wb")};
    if (!fc.f) return io_fail(PCT_E_IO, std::string(path) + ": cannot create");
    const int32_t version = 0, n = (int32_t)in->n;
    bool ok = fwrite("PCTD", 1, 4, fc.f) == 4 && fwrite(&version, 4, 1, fc.f) == 1 && fwrite(&n, 4, 1, fc.f) == 1 &&
              fwrite(&header->projection_angle, 4, 1, fc.f) == 1 && fwrite(&header->beam_energy, 4, 1, fc.f) == 1 &&
              fwrite(&header->generation_date, 4, 1, fc.f) == 1 && fwrite(&header->preprocess_date, 4, 1, fc.f) == 1;
    for (const char* s : {header->phantom_name, header->data_source, header->prepared_by}) {
        const int32_t len = (int32_t)strnlen(s, PCT_PCTD_STR - 1);
        ok = ok && fwrite(&len, 4, 1, fc.f) == 1 && (len == 0 || fwrite(s, 1, (size_t)len, fc.f) == (size_t)len);
    }
    for (const float* a : arrays)
        ok = ok && (n == 0 || fwrite(a, sizeof(float), (size_t)n, fc.f) == (size_t)n);
    if (!ok) return io_fail(PCT_E_IO, std::string(path) + ": write failed");
    if (fclose(fc.f) != 0) {
        fc.f = nullptr;
        return io_fail(PCT_E_IO, std::string(path) + ": write failed");
    }
    fc.f = nullptr;
    return PCT_OK;
}

extern "C" int pct_array_to_disk(const char* filename_base, const char* directory, const char* folder, const void* data,
                                 int type, int x_max, int y_max, int z_max, int64_t elements, int single_file) {
    if (!filename_base || !directory || !folder || (!data && elements > 0))
        return io_fail(PCT_E_INVALID, "pct_array_to_disk: null argument");
    if (x_max <= 0 || y_max <= 0 || z_max <= 0 || type < PCT_TXT_F32 || type > PCT_TXT_BOOL)
        return io_fail(PCT_E_INVALID, "pct_array_to_disk: bad dimensions or type");
    int num_files = z_max, z_start = 0, z_end = 1;
    if (single_file) {
        num_files = 1;
        z_end = z_max;
    }
    for (int file = 0; file < num_files; file++) {
        char filename[4096];
        if (num_files == z_max)
            snprintf(filename, sizeof filename, "%s%s/%s_%d.txt", directory, folder, filename_base, file);
        else
            snprintf(filename, sizeof filename, "%s%s/%s.txt", directory, folder, filename_base);
        std::ofstream out(filename);
        if (!out) return io_fail(PCT_E_IO, std::string(filename) + ": cannot create");
        switch (type) {
        case PCT_TXT_F32: dump_values(out, (const float*)data, x_max, y_max, z_start, z_end, elements); break;
        case PCT_TXT_F64: dump_values(out, (const double*)data, x_max, y_max, z_start, z_end, elements); break;
        case PCT_TXT_I32: dump_values(out, (const int32_t*)data, x_max, y_max, z_start, z_end, elements); break;
        case PCT_TXT_U32: dump_values(out, (const uint32_t*)data, x_max, y_max, z_start, z_end, elements); break;
        default: dump_values(out, (const mask_byte*)data, x_max, y_max, z_start, z_end, elements); break;
        }
        out.close();
        if (!out) return io_fail(PCT_E_IO, std::string(filename) + ": write failed");
        z_start += 1;
        z_end += 1;
    }
    return PCT_OK;
}

namespace {

/* the "%u\n" header: decimal digits, then one newline */
bool read_paths_header(FILE* f, int64_t* n) {
    int64_t v = 0;
    int c, digits = 0;
    while ((c = fgetc(f)) >= '0' && c <= '9') {
        v = v * 10 + (c - '0');
        if (++digits > 10) return false;
    }
    if (c != '\n' || digits == 0 || v > UINT32_MAX) return false;
    *n = v;
    return true;
}

}  // namespace

extern "C" int pct_mlp_paths_file_info(const char* path, int64_t* n_header, int64_t* n_paths, int64_t* total) {
    if (!path || !n_header || !n_paths || !total) return io_fail(PCT_E_INVALID, "pct_mlp_paths_file_info: null argument");
    file_closer fc{fopen(path, "rb")};
    if (!fc.f) return io_fail(PCT_E_IO, std::string(path) + ": cannot open");
    if (!read_paths_header(fc.f, n_header)) return io_fail(PCT_E_IO, std::string(path) + ": no path-count line");
    int64_t n = 0, t = 0;
    uint32_t count;
    while (fread(&count, 4, 1, fc.f) == 1) {
        if (fseek(fc.f, 4L * count, SEEK_CUR) != 0) return io_fail(PCT_E_IO, std::string(path) + ": short path");
        n += 1;
        t += count;
    }
    /* fseek past the end succeeds: check the size */
    long end = ftell(fc.f);
    if (fseek(fc.f, 0, SEEK_END) != 0 || ftell(fc.f) != end) return io_fail(PCT_E_IO, std::string(path) + ": short path");
    *n_paths = n;
    *total = t;
    return PCT_OK;
}

extern "C" int pct_mlp_paths_file_read(const char* path, int64_t max_paths, int64_t capacity, int64_t* offsets,
                                       uint32_t* voxels, int64_t* n_read) {
    if (!path || !offsets || !n_read || (capacity > 0 && !voxels))
        return io_fail(PCT_E_INVALID, "pct_mlp_paths_file_read: null argument");
    file_closer fc{fopen(path, "rb")};
    if (!fc.f) return io_fail(PCT_E_IO, std::string(path) + ": cannot open");
    int64_t header;
    if (!read_paths_header(fc.f, &header)) return io_fail(PCT_E_IO, std::string(path) + ": no path-count line");
    const int64_t want = header < max_paths ? header : max_paths;
    offsets[0] = 0;
    int64_t k = 0;
    for (; k < want; ++k) {
        uint32_t count;
        if (fread(&count, 4, 1, fc.f) != 1)
            return io_fail(PCT_E_IO, std::string(path) + ": " + std::to_string(k) + " paths, header says " +
                                         std::to_string(header));
        if (offsets[k] + (int64_t)count > capacity)
            return io_fail(PCT_E_INVALID, std::string(path) + ": intersections exceed the capacity");
        if (!read_exact(fc.f, voxels + offsets[k], 4 * (size_t)count))
            return io_fail(PCT_E_IO, std::string(path) + ": short path");
        offsets[k + 1] = offsets[k] + count;
    }
    *n_read = k;
    return PCT_OK;
}

extern "C" int pct_mlp_paths_file_write(const char* path, int64_t n, const int64_t* offsets, const uint32_t* voxels) {
    if (!path || n < 0 || n > UINT32_MAX || (n > 0 && !offsets))
        return io_fail(PCT_E_INVALID, "pct_mlp_paths_file_write: bad arguments");
    for (int64_t k = 0; k < n; ++k) {
        int64_t c = offsets[k + 1] - offsets[k];
        if (c < 0 || c > UINT32_MAX || (c > 0 && !voxels))
            return io_fail(PCT_E_INVALID, "pct_mlp_paths_file_write: bad offsets");
    }
    file_closer fc{fopen(path, "wb")};
    if (!fc.f) return io_fail(PCT_E_IO, std::string(path) + ": cannot create");
    bool ok = fprintf(fc.f, "%u\n", (unsigned)n) > 0;
    for (int64_t k = 0; ok && k < n; ++k) {
        const uint32_t count = (uint32_t)(offsets[k + 1] - offsets[k]);
        ok = fwrite(&count, 4, 1, fc.f) == 1 && (count == 0 || fwrite(voxels + offsets[k], 4, count, fc.f) == count);
    }
    int closed = fclose(fc.f);
    fc.f = nullptr;
    if (!ok || closed != 0) return io_fail(PCT_E_IO, std::string(path) + ": write failed");
    return PCT_OK;
}

extern "C" int pct_path_errors_read(const char* path, int64_t count, float* errors) {
    if (!path || count < 0 || (count > 0 && !errors)) return io_fail(PCT_E_INVALID, "pct_path_errors_read: bad arguments");
    file_closer fc{fopen(path, "rb")};
    if (!fc.f) return io_fail(PCT_E_IO, std::string(path) + ": cannot open");
    if (!read_exact(fc.f, errors, sizeof(float) * (size_t)count))
        return io_fail(PCT_E_IO, std::string(path) + ": fewer than " + std::to_string(count) + " path errors");
    return PCT_OK;
}

extern "C" int pct_path_errors_write(const char* path, int64_t count, const float* errors) {
    if (!path || count < 0 || (count > 0 && !errors)) return io_fail(PCT_E_INVALID, "pct_path_errors_write: bad arguments");
    file_closer fc{fopen(path, "wb")};
    if (!fc.f) return io_fail(PCT_E_IO, std::string(path) + ": cannot create");
    bool ok = count == 0 || fwrite(errors, sizeof(float), (size_t)count, fc.f) == (size_t)count;
    int closed = fclose(fc.f);
    fc.f = nullptr;
    if (!ok || closed != 0) return io_fail(PCT_E_IO, std::string(path) + ": write failed");
    return PCT_OK;
}
