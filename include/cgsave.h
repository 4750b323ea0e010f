/*
 * cgsave.h: the ocean's four fields in the reference's save format, format 28, on the host.
 *
 * A save is a container: "CGWD", int32 format version, int32 header length, int32 CRC-32 of the
 * header, the header (UTF-8 JSON: the document, the compression, the writer, the directory of the
 * payload's records, the payload's length), then one frame per 1 MiB chunk of the expanded payload
 * (int32 raw length, int32 stored length, int32 chunk checksum, int32 method 0 stored / 1 gzip, the
 * stored bytes) and a frame of sixteen zero bytes; little-endian throughout (WorldCodec.kt:101-116).
 * The expanded payload is a run of records, each int32 name length, the ASCII name, int32 element
 * type, int32 element count, int64 byte length and the elements: the world's lists as JSON, then
 * 56 per-cell arrays in a fixed order (WorldSections.kt:144-233), four of which are the ocean's:
 * ocean.velocityX, ocean.velocityY, ocean.temperature and ocean.anomaly, binary32 by their raw bits,
 * one per cell of the map grid, row-major (WorldSections.kt:184-187).
 *
 * These entry points need no GPU and no context. They read a whole save held in memory (a mapped
 * file will do: it is read a chunk at a time, and nothing the size of the payload is allocated),
 * and write a save whose ocean is replaced, every other byte of the payload carried as it was.
 * Everything the reference's reader holds a save to is held here too (WorldCodec.kt:525-670,
 * WorldSections.kt:656-885): the prefix, the header's checksum and version, a directory equal to the
 * format's own for the grid, every frame's lengths, method and checksum, each record's name, type,
 * count and length, the closing frame and nothing after it; and the ocean's values finite. The other
 * arrays' values and the lists' JSON are carried as bytes, held only to their records and checksums.
 *
 * Reference files are cited relative to cartography/src/commonMain/kotlin/com/cartogenesis/cartography/.
 */
#ifndef CGSAVE_H
#define CGSAVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* What a refused save is: SaveProblem (SaveRefusal.kt:11-45). CG_OK (0) is a save that opens; a
 * bad argument (a null pointer, a buffer too small) is CG_ERR_ARG (-1) of cgocean.h. */
#define CG_SAVE_NOT_A_SAVE (-20)
#define CG_SAVE_INCOMPLETE (-21)
#define CG_SAVE_DAMAGED (-22)
#define CG_SAVE_WRONG_VERSION (-23)
#define CG_SAVE_TOO_LARGE (-24)
#define CG_SAVE_CANNOT_EXPAND (-25)

#define CG_SAVE_FORMAT_VERSION 28           /* WorldCodec.FORMAT_VERSION (WorldCodec.kt:338) */
#define CG_SAVE_CHUNK_BYTES (1 << 20)       /* WorldCodec.CHUNK_BYTES (WorldCodec.kt:364) */
#define CG_SAVE_LARGEST_GRID_CELLS (2L * 4096 * 4096) /* WorldCodec.LARGEST_GRID_CELLS */
#define CG_SAVE_SECTIONS 57                 /* the directory's entries: the lists and 56 arrays */

/* SectionType (WorldSections.kt:28-45): the element type codes of a record. */
#define CG_SECTION_F32 1
#define CG_SECTION_I32 2
#define CG_SECTION_U8 3
#define CG_SECTION_UTF8 4

/* SectionInfo (WorldSections.kt:54-67): a directory entry. `offset` is where the record starts in
 * the expanded payload; `bytes` the elements' own bytes. */
typedef struct {
    char name[48];
    int type;       /* CG_SECTION_* */
    int count;      /* elements: cells, or bytes of UTF-8 for the lists */
    int64_t bytes;
    int64_t offset;
} cg_section_info;

/* Crc32.of (PngWriter.kt:378-407): the reflected CRC-32 of `len` bytes, run on from the checksum
 * `continuing` of bytes before them (0 starts afresh). */
uint32_t cg_save_crc32(const uint8_t *data, size_t len, uint32_t continuing);

/* chunkChecksum (WorldSections.kt:496-499): the CRC-32 of a chunk's raw bytes run on from the
 * header's checksum and the chunk's index as int32. */
uint32_t cg_save_chunk_checksum(uint32_t header_checksum, int index, const uint8_t *raw, size_t len);

/* WorldSections.directory (WorldSections.kt:251-265): the directory of a save of `cells` cells
 * whose lists are `lists_bytes` bytes of JSON, into out[0 .. min(cap, CG_SAVE_SECTIONS)); returns
 * CG_SAVE_SECTIONS, or CG_ERR_ARG. *payload_bytes (may be NULL) = WorldSections.payloadBytes. */
int cg_save_directory(int cells, int lists_bytes, cg_section_info *out, int cap, int64_t *payload_bytes);

/* The four ocean records (ocean.velocityX, ocean.velocityY, ocean.temperature, ocean.anomaly), as
 * they stand, one after the other, in the expanded payload of every save of `cells` cells: what
 * PayloadWriter.record and floats write for them (WorldSections.kt:268-279, 526-578). Needs
 * 4 x (20 + name + 4 cells) bytes; *written = that. Refuses (CG_ERR_ARG) a value that is not finite,
 * which the reader would refuse (WorldSections.kt:101-102, 744-747). */
int cg_ocean_records_write(int cells, const float *velocity_x, const float *velocity_y, const float *temperature,
                           const float *anomaly, uint8_t *out, size_t cap, size_t *written);

/* The reverse: each record held to its directory entry (PayloadReader.expectRecord,
 * WorldSections.kt:676-690) and each value to finite. `records` starts at ocean.velocityX. */
int cg_ocean_records_read(const uint8_t *records, size_t len, int cells, float *velocity_x, float *velocity_y,
                          float *temperature, float *anomaly, char *reason, size_t reason_len);

/* The grid of a save, from its prefix and header alone (WorldCodec.decodeHeader, WorldCodec.kt:505-517):
 * the columns and rows of its map grid, the expanded payload's length, and the most bytes a save
 * with its ocean replaced can take (every chunk stored). Any pointer may be NULL. */
int cg_save_inspect(const uint8_t *file, size_t len, int *across, int *down, int64_t *payload_bytes,
                    int64_t *largest_rewrite, char *reason, size_t reason_len);

/* WorldCodec.read (WorldCodec.kt:525-546) for the ocean: the whole save checked, from its first byte
 * to its last, and its four ocean arrays (across x down floats each) copied out. */
int cg_save_read_ocean(const uint8_t *file, size_t len, float *velocity_x, float *velocity_y, float *temperature,
                       float *anomaly, char *reason, size_t reason_len);

/* The save `file` with its ocean replaced: the same prefix and header (so the same document and
 * checksums), the same chunks with the four arrays' elements written over, each chunk framed again
 * (PayloadWriter.flush, WorldSections.kt:627-640): compressed with gzip when the header names gzip,
 * `store` is 0 and that is smaller, stored otherwise. The input is checked as by cg_save_read_ocean
 * first. out needs at most cg_save_inspect's largest_rewrite bytes; *written = the save's length. */
int cg_save_replace_ocean(const uint8_t *file, size_t len, const float *velocity_x, const float *velocity_y,
                          const float *temperature, const float *anomaly, int store, uint8_t *out, size_t cap,
                          size_t *written, char *reason, size_t reason_len);

#ifdef __cplusplus
}
#endif

#endif /* CGSAVE_H */
