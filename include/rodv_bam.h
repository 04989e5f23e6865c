/*
 * rodv_bam.h -- BAM input for the PairHMM engine: a BGZF/BAM reader and
 * writer with BAI region queries, written over zlib (no htslib), and the
 * per-read step that turns a record into PairHMM read inputs.
 *
 * It replaces the reference's BamHandle (src/iostream/bam_handle.h:6-88,
 * bam_handle.cpp:1-75; the query types of genomic_interval_handle.h:9-13):
 *   BamHandle(path, READ)            -> rodv_bam_open (loads path.bai if present)
 *   BamHandle(path, WRITE, level)    -> rodv_bam_create
 *   next(record) / has_next()        -> rodv_bam_next (1: a record, 0: the end)
 *   query_from_region(region)        -> rodv_bam_query_region ("chr", "chr:beg", "chr:beg-end")
 *   query_from_region(tid, beg, end) -> rodv_bam_query (0-based, end exclusive)
 *   query_from_multi_region(r, n)    -> rodv_bam_query_multi
 *   write_record / set_header        -> rodv_bam_write / the header arguments of rodv_bam_create
 *   close()                          -> rodv_bam_close (a writer also writes the BGZF end
 *                                       block and, for coordinate-sorted output, path.bai)
 * Errors are return codes (the reference throws std::runtime_error); the
 * record fields follow the BAM format (SAM specification, section 4).
 * The reference has no pipeline that feeds BAM reads to the PairHMM (its
 * rodv_main is empty) and no BAM fixture (tests/test_bam_reader.cpp is an
 * empty file).
 */
#ifndef RODV_BAM_H_
#define RODV_BAM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RODV_BAM_OK = 0,
    RODV_BAM_ERR_IO = -1,        /* open, read, write or seek failed        */
    RODV_BAM_ERR_FORMAT = -2,    /* not BGZF/BAM/BAI, or a corrupt block    */
    RODV_BAM_ERR_ARG = -3,       /* invalid argument or mode                */
    RODV_BAM_ERR_NO_INDEX = -4,  /* a region query without a BAI index      */
    RODV_BAM_ERR_REGION = -5     /* unknown reference name or bad interval  */
};

typedef struct rodv_bam rodv_bam;

/* One alignment record.  For rodv_bam_next the pointers stay valid until the
 * next call on the handle.  seq holds the bases as letters of
 * "=ACMGRSVTWYHKDBN"; qual the Phred values (0xFF: absent); aux the raw
 * auxiliary fields. */
typedef struct rodv_bam_record {
    int32_t tid, pos;            /* reference index and 0-based leftmost position (-1: none) */
    int32_t end;                 /* reference end (exclusive) from the CIGAR; pos + 1 if it
                                    consumes no reference                               */
    uint32_t mapq, flag;
    int32_t next_tid, next_pos, tlen;
    uint32_t l_seq, n_cigar, l_aux;
    const char *name;            /* NUL-terminated                                     */
    const uint32_t *cigar;       /* op length << 4 | op (MIDNSHP=X -> 0..8)            */
    const uint8_t *seq;
    const uint8_t *qual;
    const uint8_t *aux;
} rodv_bam_record;

int rodv_bam_open(const char *path, rodv_bam **out);
int rodv_bam_load_index(rodv_bam *b, const char *bai_path); /* NULL: path + ".bai"        */
int rodv_bam_num_targets(const rodv_bam *b);
int rodv_bam_target(const rodv_bam *b, int32_t tid, const char **name, int64_t *len);
const char *rodv_bam_header_text(const rodv_bam *b);
int rodv_bam_next(rodv_bam *b, rodv_bam_record *rec);
int rodv_bam_query(rodv_bam *b, int32_t tid, int64_t beg, int64_t end);
int rodv_bam_query_region(rodv_bam *b, const char *region);
int rodv_bam_query_multi(rodv_bam *b, const char *const *regions, int num_regions);

/* Writer: header text and targets, BGZF compression level 0-9 (-1: zlib's
 * default). */
int rodv_bam_create(const char *path, const char *header_text, int32_t num_targets,
                    const char *const *target_names, const int64_t *target_lens, int level,
                    rodv_bam **out);
/* rec->end and rec->n_cigar/cigar give the bin; seq letters outside
 * "=ACMGRSVTWYHKDBN" are written as N. */
int rodv_bam_write(rodv_bam *b, const rodv_bam_record *rec);
int rodv_bam_close(rodv_bam *b);

/* The value of the Z-type auxiliary field `tag` (NULL if absent). */
const char *rodv_bam_aux_z(const rodv_bam_record *rec, const char tag[2]);

/*
 * PairHMM read inputs of a record: its bases (A, C, G, T; every other code
 * as N), base qualities, and insertion / deletion qualities from the BI / BD
 * fields (Phred + 33 strings, as GATK's base recalibration writes them) or
 * default_ins / default_del, gap-continuation qualities default_gcp.
 * clip_soft = 1 drops the soft-clipped bases at either end (CIGAR S).  A
 * record without qualities (0xFF) gets default_q.  The buffers hold
 * rec->l_seq bytes; *len receives the bases written.
 */
int rodv_bam_read_inputs(const rodv_bam_record *rec, int clip_soft, uint8_t default_q,
                         uint8_t default_ins, uint8_t default_del, uint8_t default_gcp,
                         uint8_t *bases, uint8_t *q, uint8_t *ins, uint8_t *del, uint8_t *gcp,
                         uint32_t *len);

#ifdef __cplusplus
}
#endif
#endif
