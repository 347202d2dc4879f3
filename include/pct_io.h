/*
 * pct_io.h -- C ABI of the scan files and text dumps around the path (host
 * code in libpct.so; no GPU work, no context needed).
 *
 *   PCTD version-0 projection files: one file per gantry angle, named
 *   "<directory><folder>/<base>_<angle %03d><extension>" (.cu:1011, 1782),
 *   read by count_histories_v0 (.cu:997-1097) and read_data_chunk_v02
 *   (.cu:1769-1933). Layout (little endian): the 4 bytes "PCTD", version
 *   (int32), number of events N (int32), projection angle (float, degrees),
 *   beam energy (float, MeV), generation date and pre-process date (int32,
 *   Unix time), then three strings (phantom name, data source, prepared by),
 *   each an int32 length followed by that many bytes; then 13 float arrays of
 *   N values each: t0 t1 t2 t3, v0 v1 v2 v3, u0 u1 u2 u3 (tracker planes:
 *   entry pair, exit pair) and WEPL, in mm when the scan is in mm.
 *
 *   Text dumps of images, masks and sinograms: array_2_disk (.cu:8092-8134).
 *
 * Every function returns PCT_OK or a negative PCT_E_* code; pct_io_last_error
 * gives the text of the last error of the calling thread.
 */
#ifndef PCT_IO_H
#define PCT_IO_H

#include <stdint.h>
#include "pct.h"
#include "pct_synth_api.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PCT_E_IO          (-5)   /* file missing, unreadable or malformed      */
#define PCT_PCTD_STR      256    /* stored length of the header's strings      */

typedef struct pct_pctd_header {
    int32_t version;             /* format version (0 is read and written)       */
    int32_t n_histories;         /* events in the file                           */
    float   projection_angle;    /* [deg]                                        */
    float   beam_energy;         /* [MeV]                                        */
    int32_t generation_date;     /* Unix time                                    */
    int32_t preprocess_date;     /* Unix time                                    */
    char    phantom_name[PCT_PCTD_STR];   /* NUL-terminated, cut at 255 bytes    */
    char    data_source[PCT_PCTD_STR];
    char    prepared_by[PCT_PCTD_STR];
    int64_t data_offset;         /* byte offset of the first event array         */
} pct_pctd_header;

/* The header of one projection file: replaces the per-file part of
 * count_histories_v0 (.cu:1011-1060). Versions other than 0 are refused. */
int pct_pctd_read_header(const char* path, pct_pctd_header* header);

/* The events of one projection file into host arrays of at least `capacity`
 * elements (out->gantry_angle included): replaces the version-0 branch of
 * read_data_chunk_v02 (.cu:1816-1928). With data_in_mm the 12 coordinates and
 * the WEPL are converted to cm as the reference does (the float times the
 * double 0.1, rounded to float); every event gets the gantry angle
 * (int(projection angle) + 270) % 360. *n_read = N. */
int pct_pctd_read(const char* path, int data_in_mm, int64_t capacity, const pct_synth_soa* out, int64_t* n_read);

/* Writes a version-0 projection file with the header's fields (n_histories
 * is taken from in->n) and the 13 arrays of `in` as they are (no unit
 * conversion; gantry_angle is not stored, the projection angle is). No
 * counterpart in the reference, which only reads; it lets synthetic scans
 * be stored in the reference's input format. */
int pct_pctd_write(const char* path, const pct_pctd_header* header, const pct_raw_soa* in);

/* Element types of pct_array_to_disk */
#define PCT_TXT_F32  0
#define PCT_TXT_F64  1
#define PCT_TXT_I32  2
#define PCT_TXT_U32  3
#define PCT_TXT_BOOL 4           /* one byte per element, printed 0 / 1          */

/* array_2_disk (.cu:8092-8134): the first `elements` values of data, x
 * fastest, written as text with the C++ stream's default formatting, each
 * value followed by a space and each row of x_max values by a newline; one
 * file "<directory><folder>/<base>.txt" holding every z (single_file), or one
 * file "<directory><folder>/<base>_<z>.txt" per z. */
int pct_array_to_disk(const char* filename_base, const char* directory, const char* folder, const void* data,
                      int type, int x_max, int y_max, int z_max, int64_t elements, int single_file);

/* MLP path files of the research loop (.cu:7479-7493, 4746-4761): the
 * number of paths as text ("%u\n"), then per path its intersection count
 * (uint32) and that many voxel indices (uint32). Path-error files hold one
 * float per intersection, path after path, with no header (.cu:4758-4761).
 * The header is read up to its one newline; the reference's fscanf("%u\n")
 * would also skip a first count whose low byte is a whitespace code. */
/* Number of paths the header announces, paths present, and their total
 * intersection count. */
int pct_mlp_paths_file_info(const char* path, int64_t* n_header, int64_t* n_paths, int64_t* total);
/* Up to max_paths paths (the reference reads min(header count,
 * reconstruction histories), .cu:7522) into host arrays: offsets (n+1,
 * offsets[0] = 0) and voxels (at least `capacity` values). */
int pct_mlp_paths_file_read(const char* path, int64_t max_paths, int64_t capacity, int64_t* offsets, uint32_t* voxels,
                            int64_t* n_read);
/* write_MLP_path over n lists, with the header (.cu:7480-7491) */
int pct_mlp_paths_file_write(const char* path, int64_t n, const int64_t* offsets, const uint32_t* voxels);
/* read_MLP_path_error (.cu:4758) for `count` intersections, and its writer
 * (no counterpart in the reference: its error files come from elsewhere) */
int pct_path_errors_read(const char* path, int64_t count, float* errors);
int pct_path_errors_write(const char* path, int64_t count, const float* errors);

const char* pct_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
