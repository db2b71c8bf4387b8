/*
 * unifrac-binaries_amd -- C ABI of libssu.so, the MI355X (gfx950) Strided
 * State UniFrac / Faith PD library.
 *
 * Drop-in contract: the type layouts, enum values and entry points below are
 * those of biocore/unifrac-binaries src/api.hpp and src/status_enum.hpp, so a
 * caller built against the reference header links against this library
 * unchanged.  Each declaration names the reference declaration it replaces.
 * Declarations after "Extensions" have no counterpart in the reference.
 */
#ifndef SSU_AMD_H
#define SSU_AMD_H

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
#define SSU_EXTERN extern "C"
#else
#define SSU_EXTERN
#endif

/* src/api.hpp:22-23 */
#define PARTIAL_MAGIC "SSU-PARTIAL-01"
#define PARTIAL_MAGIC_V2 0x088ABA02

/* src/status_enum.hpp:4-6 -- numeric values are part of the ABI */
typedef enum compute_status {okay = 0, tree_missing, table_missing, table_empty, unknown_method,
                             table_and_tree_do_not_overlap, output_error, invalid_method, grouping_missing} ComputeStatus;
typedef enum io_status {read_okay = 0, write_okay, open_error, read_error, magic_incompatible, bad_header,
                        unexpected_end, write_error} IOStatus;
typedef enum merge_status {merge_okay = 0, incomplete_stripe_set, sample_id_consistency, square_mismatch,
                           partials_mismatch, stripes_overlap} MergeStatus;

/* src/api.hpp:40-46: condensed upper triangle, scipy squareform order */
typedef struct mat {
    unsigned int n_samples;
    unsigned int cf_size;
    bool is_upper_triangle;
    double* condensed_form;
    char** sample_ids;
} mat_t;

/* src/api.hpp:54-59 */
typedef struct mat_full_fp64 {
    uint32_t n_samples;
    uint32_t flags;   /* 0, or (fd & 0x0fff) | 0x1000 when the matrix is an mmap'd file */
    double* matrix;
    char** sample_ids;
} mat_full_fp64_t;

/* src/api.hpp:67-72 */
typedef struct mat_full_fp32 {
    uint32_t n_samples;
    uint32_t flags;
    float* matrix;
    char** sample_ids;
} mat_full_fp32_t;

/* src/api.hpp:82-86 */
typedef struct results_vec {
    unsigned int n_samples;
    double* values;
    char** sample_ids;
} r_vec;

/* src/api.hpp:99-107: stripes[s - stripe_start][k] = d(k, (k + s + 1) mod n_samples) */
typedef struct partial_mat {
    uint32_t n_samples;
    char** sample_ids;
    double** stripes;
    uint32_t stripe_start;
    uint32_t stripe_stop;
    uint32_t stripe_total;
    bool is_upper_triangle;
} partial_mat_t;

/* src/api.hpp:122-132 */
typedef struct partial_dyn_mat {
    uint32_t n_samples;
    char** sample_ids;
    uint64_t* offsets;
    double** stripes;
    uint32_t stripe_start;
    uint32_t stripe_stop;
    uint32_t stripe_total;
    bool is_upper_triangle;
    char* filename;
} partial_dyn_mat_t;

/* src/api.hpp:145-154: observation-major CSR */
typedef struct support_biom {
    char** obs_ids;
    char** sample_ids;
    uint32_t* indices;
    uint32_t* indptr;
    double* data;
    int n_obs;
    int n_samples;
    int nnz;
} support_biom_t;

/* src/api.hpp:163-168: balanced parentheses, per-parenthesis lengths and names */
typedef struct support_bptree {
    bool* structure;
    double* lengths;
    char** names;
    int n_parens;
} support_bptree_t;

/* src/api.hpp:173-175 */
typedef struct opaque_bptree {
    int dummy;
} opaque_bptree_t;

/* ------------------------------------------------------------------ trees */
SSU_EXTERN void ssu_set_random_seed(unsigned int new_seed);                                          /* api.hpp:28 */
SSU_EXTERN IOStatus read_bptree_opaque(const char* tree_filename, opaque_bptree_t** tree_data);      /* api.hpp:188 */
SSU_EXTERN void load_bptree_opaque(const char* newick, opaque_bptree_t** tree_data);                /* api.hpp:191 */
SSU_EXTERN void convert_bptree_opaque(const support_bptree_t* in_tree, opaque_bptree_t** tree_data); /* api.hpp:194 */
SSU_EXTERN void destroy_bptree_opaque(opaque_bptree_t** tree_data);                                  /* api.hpp:203 */
SSU_EXTERN int get_bptree_opaque_els(opaque_bptree_t* tree_data);                                    /* api.hpp:206 */

/* ------------------------------------------------------------- destructors */
SSU_EXTERN void destroy_mat(mat_t** result);                          /* api.hpp:196 */
SSU_EXTERN void destroy_mat_full_fp64(mat_full_fp64_t** result);      /* api.hpp:197 */
SSU_EXTERN void destroy_mat_full_fp32(mat_full_fp32_t** result);      /* api.hpp:198 */
SSU_EXTERN void destroy_partial_mat(partial_mat_t** result);          /* api.hpp:199 */
SSU_EXTERN void destroy_partial_dyn_mat(partial_dyn_mat_t** result);  /* api.hpp:200 */
SSU_EXTERN void destroy_results_vec(r_vec** result);                  /* api.hpp:201 */

/* ------------------------------------------------------ condensed, from files */
SSU_EXTERN ComputeStatus one_off_v3(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                                    bool variance_adjust, double alpha, bool bypass_tips, bool normalize_sample_counts,
                                    unsigned int n_substeps, mat_t** result);                        /* api.hpp:228 */
SSU_EXTERN ComputeStatus one_off_wtree_v3(const char* biom_filename, const opaque_bptree_t* tree_data,
                                          const char* unifrac_method, bool variance_adjust, double alpha,
                                          bool bypass_tips, bool normalize_sample_counts, unsigned int n_substeps,
                                          mat_t** result);                                           /* api.hpp:233 */
SSU_EXTERN ComputeStatus one_off(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                                 bool variance_adjust, double alpha, bool bypass_tips, unsigned int n_substeps,
                                 mat_t** result);                                                    /* api.hpp:238 */
SSU_EXTERN ComputeStatus one_off_wtree(const char* biom_filename, const opaque_bptree_t* tree_data,
                                       const char* unifrac_method, bool variance_adjust, double alpha,
                                       bool bypass_tips, unsigned int n_substeps, mat_t** result);  /* api.hpp:242 */

/* ----------------------------------------------------- full matrix, in memory */
SSU_EXTERN ComputeStatus one_off_matrix_inmem_v3(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                                 const char* unifrac_method, bool variance_adjust, double alpha,
                                                 bool bypass_tips, bool normalize_sample_counts,
                                                 unsigned int n_substeps, unsigned int subsample_depth,
                                                 bool subsample_with_replacement, const char* mmap_dir,
                                                 mat_full_fp64_t** result);                          /* api.hpp:268 */
SSU_EXTERN ComputeStatus one_off_matrix_inmem_v2(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                                 const char* unifrac_method, bool variance_adjust, double alpha,
                                                 bool bypass_tips, unsigned int n_substeps,
                                                 unsigned int subsample_depth, bool subsample_with_replacement,
                                                 const char* mmap_dir, mat_full_fp64_t** result);    /* api.hpp:275 */
SSU_EXTERN ComputeStatus one_off_inmem(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                       const char* unifrac_method, bool variance_adjust, double alpha,
                                       bool bypass_tips, unsigned int n_substeps, mat_full_fp64_t** result); /* api.hpp:282 */
SSU_EXTERN ComputeStatus one_off_matrix_inmem_fp32_v3(const support_biom_t* table_data,
                                                      const support_bptree_t* tree_data, const char* unifrac_method,
                                                      bool variance_adjust, double alpha, bool bypass_tips,
                                                      bool normalize_sample_counts, unsigned int n_substeps,
                                                      unsigned int subsample_depth, bool subsample_with_replacement,
                                                      const char* mmap_dir, mat_full_fp32_t** result); /* api.hpp:307 */
SSU_EXTERN ComputeStatus one_off_matrix_inmem_fp32_v2(const support_biom_t* table_data,
                                                      const support_bptree_t* tree_data, const char* unifrac_method,
                                                      bool variance_adjust, double alpha, bool bypass_tips,
                                                      unsigned int n_substeps, unsigned int subsample_depth,
                                                      bool subsample_with_replacement, const char* mmap_dir,
                                                      mat_full_fp32_t** result);                     /* api.hpp:314 */
SSU_EXTERN ComputeStatus one_off_inmem_fp32(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                            const char* unifrac_method, bool variance_adjust, double alpha,
                                            bool bypass_tips, unsigned int n_substeps,
                                            mat_full_fp32_t** result);                               /* api.hpp:321 */

/* -------------------------------------------------------- full matrix, files */
SSU_EXTERN ComputeStatus one_off_matrix_v3(const char* biom_filename, const char* tree_filename,
                                           const char* unifrac_method, bool variance_adjust, double alpha,
                                           bool bypass_tips, bool normalize_sample_counts, unsigned int n_substeps,
                                           unsigned int subsample_depth, bool subsample_with_replacement,
                                           const char* mmap_dir, mat_full_fp64_t** result);          /* api.hpp:348 */
SSU_EXTERN ComputeStatus one_off_matrix_v3t(const char* biom_filename, const opaque_bptree_t* tree_data,
                                            const char* unifrac_method, bool variance_adjust, double alpha,
                                            bool bypass_tips, bool normalize_sample_counts, unsigned int n_substeps,
                                            unsigned int subsample_depth, bool subsample_with_replacement,
                                            const char* mmap_dir, mat_full_fp64_t** result);         /* api.hpp:355 */
SSU_EXTERN ComputeStatus one_off_matrix_v2(const char* biom_filename, const char* tree_filename,
                                           const char* unifrac_method, bool variance_adjust, double alpha,
                                           bool bypass_tips, unsigned int n_substeps, unsigned int subsample_depth,
                                           bool subsample_with_replacement, const char* mmap_dir,
                                           mat_full_fp64_t** result);                                /* api.hpp:362 */
SSU_EXTERN ComputeStatus one_off_matrix_v2t(const char* biom_filename, const opaque_bptree_t* tree_data,
                                            const char* unifrac_method, bool variance_adjust, double alpha,
                                            bool bypass_tips, unsigned int n_substeps, unsigned int subsample_depth,
                                            bool subsample_with_replacement, const char* mmap_dir,
                                            mat_full_fp64_t** result);                               /* api.hpp:369 */
SSU_EXTERN ComputeStatus one_off_matrix(const char* biom_filename, const char* tree_filename,
                                        const char* unifrac_method, bool variance_adjust, double alpha,
                                        bool bypass_tips, unsigned int n_substeps, const char* mmap_dir,
                                        mat_full_fp64_t** result);                                   /* api.hpp:375 */
SSU_EXTERN ComputeStatus one_off_matrix_fp32_v3(const char* biom_filename, const char* tree_filename,
                                                const char* unifrac_method, bool variance_adjust, double alpha,
                                                bool bypass_tips, bool normalize_sample_counts,
                                                unsigned int n_substeps, unsigned int subsample_depth,
                                                bool subsample_with_replacement, const char* mmap_dir,
                                                mat_full_fp32_t** result);                           /* api.hpp:404 */
SSU_EXTERN ComputeStatus one_off_matrix_fp32_v3t(const char* biom_filename, const opaque_bptree_t* tree_data,
                                                 const char* unifrac_method, bool variance_adjust, double alpha,
                                                 bool bypass_tips, bool normalize_sample_counts,
                                                 unsigned int n_substeps, unsigned int subsample_depth,
                                                 bool subsample_with_replacement, const char* mmap_dir,
                                                 mat_full_fp32_t** result);                          /* api.hpp:411 */
SSU_EXTERN ComputeStatus one_off_matrix_fp32_v2(const char* biom_filename, const char* tree_filename,
                                                const char* unifrac_method, bool variance_adjust, double alpha,
                                                bool bypass_tips, unsigned int n_substeps,
                                                unsigned int subsample_depth, bool subsample_with_replacement,
                                                const char* mmap_dir, mat_full_fp32_t** result);     /* api.hpp:418 */
SSU_EXTERN ComputeStatus one_off_matrix_fp32_v2t(const char* biom_filename, const opaque_bptree_t* tree_data,
                                                 const char* unifrac_method, bool variance_adjust, double alpha,
                                                 bool bypass_tips, unsigned int n_substeps,
                                                 unsigned int subsample_depth, bool subsample_with_replacement,
                                                 const char* mmap_dir, mat_full_fp32_t** result);    /* api.hpp:425 */
SSU_EXTERN ComputeStatus one_off_matrix_fp32(const char* biom_filename, const char* tree_filename,
                                             const char* unifrac_method, bool variance_adjust, double alpha,
                                             bool bypass_tips, unsigned int n_substeps, const char* mmap_dir,
                                             mat_full_fp32_t** result);                              /* api.hpp:432 */

/* --------------------------------------------------------------- dense pair */
SSU_EXTERN ComputeStatus one_dense_pair_v3t(unsigned int n_obs, const char** obs_ids, const double* sample1,
                                            const double* sample2, const opaque_bptree_t* tree_data,
                                            const char* unifrac_method, bool variance_adjust, double alpha,
                                            bool bypass_tips, bool normalize_sample_counts, double* result); /* api.hpp:456 */
SSU_EXTERN ComputeStatus one_dense_pair_v3(unsigned int n_obs, const char** obs_ids, const double* sample1,
                                           const double* sample2, const support_bptree_t* tree_data,
                                           const char* unifrac_method, bool variance_adjust, double alpha,
                                           bool bypass_tips, bool normalize_sample_counts, double* result);  /* api.hpp:462 */
SSU_EXTERN ComputeStatus one_dense_pair_v2t(unsigned int n_obs, const char** obs_ids, const double* sample1,
                                            const double* sample2, const opaque_bptree_t* tree_data,
                                            const char* unifrac_method, bool variance_adjust, double alpha,
                                            bool bypass_tips, double* result);                       /* api.hpp:469 */
SSU_EXTERN ComputeStatus one_dense_pair_v2(unsigned int n_obs, const char** obs_ids, const double* sample1,
                                           const double* sample2, const support_bptree_t* tree_data,
                                           const char* unifrac_method, bool variance_adjust, double alpha,
                                           bool bypass_tips, double* result);                        /* api.hpp:475 */

/* ----------------------------------------------------------------- Faith PD */
SSU_EXTERN ComputeStatus faith_pd_one_off(const char* biom_filename, const char* tree_filename,
                                          r_vec** result);                                           /* api.hpp:492 */

/* ------------------------------------------------------------------ partial */
SSU_EXTERN ComputeStatus partial_v3(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                                    bool variance_adjust, double alpha, bool bypass_tips, bool normalize_sample_count,
                                    unsigned int n_substeps, unsigned int stripe_start, unsigned int stripe_stop,
                                    partial_mat_t** result);                                         /* api.hpp:805 */
SSU_EXTERN ComputeStatus partial(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                                 bool variance_adjust, double alpha, bool bypass_tips, unsigned int n_substeps,
                                 unsigned int stripe_start, unsigned int stripe_stop,
                                 partial_mat_t** result);                                            /* api.hpp:811 */

/* ------------------------------------------------------------ text writers */
SSU_EXTERN IOStatus write_mat(const char* filename, mat_t* result);                              /* api.hpp:649 */
SSU_EXTERN IOStatus write_mat_from_matrix(const char* filename, mat_full_fp64_t* result);        /* api.hpp:688 */
SSU_EXTERN IOStatus write_mat_from_matrix_fp32(const char* filename, mat_full_fp32_t* result);   /* api.hpp:691 */
SSU_EXTERN IOStatus write_vec(const char* filename, r_vec* result);                              /* api.hpp:763 */

/* Distance-matrix HDF5 files (format "BDSM", version "2020.12"): "order", and
 * "matrix" when save_dist; pcoa_dims > 0 adds "pcoa_method" ("FSVD"),
 * "pcoa_eigvals", "pcoa_proportion_explained" and "pcoa_samples". */
SSU_EXTERN IOStatus write_mat_from_matrix_hdf5_fp64(const char* filename, mat_full_fp64_t* result,
                                                    unsigned int pcoa_dims, int save_dist);  /* api.hpp:722 */
SSU_EXTERN IOStatus write_mat_from_matrix_hdf5_fp32(const char* filename, mat_full_fp32_t* result,
                                                    unsigned int pcoa_dims, int save_dist);  /* api.hpp:752 */
/* As above plus the statistics datasets (stat_methods, stat_test_names,
 * stat_grouping_names, stat_n_groups, stat_values, stat_pvalues,
 * stat_n_permutations) when stat_n_vals > 0. */
SSU_EXTERN IOStatus write_mat_from_matrix_hdf5_fp64_v2(const char* output_filename, mat_full_fp64_t* result,
                                                       unsigned int pcoa_dims, int save_dist, unsigned int stat_n_vals,
                                                       const char** stat_method_arr, const char** stat_name_arr,
                                                       const double* stat_val_arr, const double* stat_pval_arr,
                                                       const unsigned int* stat_perm_count_arr,
                                                       const char** stat_group_name_arr,
                                                       const unsigned int* stat_group_count_arr);   /* api.hpp:714 */
SSU_EXTERN IOStatus write_mat_from_matrix_hdf5_fp32_v2(const char* output_filename, mat_full_fp32_t* result,
                                                       unsigned int pcoa_dims, int save_dist, unsigned int stat_n_vals,
                                                       const char** stat_method_arr, const char** stat_name_arr,
                                                       const float* stat_val_arr, const float* stat_pval_arr,
                                                       const unsigned int* stat_perm_count_arr,
                                                       const char** stat_group_name_arr,
                                                       const unsigned int* stat_group_count_arr);   /* api.hpp:744 */

/* PERMANOVA of each grouping column (TSV file: header line, sample id in the
 * first field) on the matrix: pseudo-F, permutation p-value and number of
 * groups per column; grouping_missing when a column or a sample is absent. */
SSU_EXTERN ComputeStatus compute_permanova_fp64(const char* grouping_filename, unsigned int n_columns,
                                                const char** columns, mat_full_fp64_t* result,
                                                unsigned int permanova_perms, double* fstats, double* pvalues,
                                                unsigned int* n_groups);                            /* api.hpp:616 */
SSU_EXTERN ComputeStatus compute_permanova_fp32(const char* grouping_filename, unsigned int n_columns,
                                                const char** columns, mat_full_fp32_t* result,
                                                unsigned int permanova_perms, float* fstats, float* pvalues,
                                                unsigned int* n_groups);                            /* api.hpp:636 */

/* One matrix from files straight to an output file.  format: hdf5 (precision
 * of the method), hdf5_fp32, hdf5_fp64 or hdf5_nodist.  pcoa_dims > 0 adds
 * the PCoA datasets, permanova_perms > 0 the PERMANOVA statistics of each
 * comma-separated grouping column. */
SSU_EXTERN ComputeStatus unifrac_to_file_v3(const char* biom_filename, const char* tree_filename,
                                            const char* out_filename, const char* unifrac_method,
                                            bool variance_adjust, double alpha, bool bypass_tips,
                                            bool normalize_sample_counts, unsigned int n_substeps, const char* format,
                                            unsigned int subsample_depth, bool subsample_with_replacement,
                                            unsigned int pcoa_dims, unsigned int permanova_perms,
                                            const char* grouping_filename, const char* grouping_columns,
                                            const char* mmap_dir);                                  /* api.hpp:524 */
SSU_EXTERN ComputeStatus unifrac_to_txt_file_v3(const char* biom_filename, const char* tree_filename,
                                                const char* out_filename, const char* unifrac_method,
                                                bool variance_adjust, double alpha, bool bypass_tips,
                                                bool normalize_sample_counts, unsigned int n_substeps,
                                                const char* mmap_dir);                              /* api.hpp:533 */
SSU_EXTERN ComputeStatus unifrac_to_file_v2(const char* biom_filename, const char* tree_filename,
                                            const char* out_filename, const char* unifrac_method,
                                            bool variance_adjust, double alpha, bool bypass_tips,
                                            unsigned int n_substeps, const char* format, unsigned int subsample_depth,
                                            bool subsample_with_replacement, unsigned int pcoa_dims,
                                            unsigned int permanova_perms, const char* grouping_filename,
                                            const char* grouping_columns, const char* mmap_dir);    /* api.hpp:539 */
SSU_EXTERN ComputeStatus unifrac_to_file(const char* biom_filename, const char* tree_filename,
                                         const char* out_filename, const char* unifrac_method, bool variance_adjust,
                                         double alpha, bool bypass_tips, unsigned int n_substeps, const char* format,
                                         unsigned int pcoa_dims, const char* mmap_dir);             /* api.hpp:548 */

/* Rarefaction replicates to one HDF5 file (format "SUJK", version "2023.03"):
 * "order" and, unless format is hdf5_nodist, "matrix:<i>" for i < n_subsamples,
 * "pcoa_*:<i>" when pcoa_dims > 0 and the stat_* datasets when
 * permanova_perms > 0.  The replicates are drawn and computed on the device,
 * from one HIP graph per batch of replicates. */
SSU_EXTERN ComputeStatus unifrac_multi_to_file_v3(const char* biom_filename, const char* tree_filename,
                                                  const char* out_filename, const char* unifrac_method,
                                                  bool variance_adjust, double alpha, bool bypass_tips,
                                                  bool normalize_sample_counts, unsigned int n_substeps,
                                                  const char* format, unsigned int n_subsamples,
                                                  unsigned int subsample_depth, bool subsample_with_replacement,
                                                  unsigned int pcoa_dims, unsigned int permanova_perms,
                                                  const char* grouping_filename, const char* grouping_columns,
                                                  const char* mmap_dir);                            /* api.hpp:583 */
SSU_EXTERN ComputeStatus unifrac_multi_to_file_v2(const char* biom_filename, const char* tree_filename,
                                                  const char* out_filename, const char* unifrac_method,
                                                  bool variance_adjust, double alpha, bool bypass_tips,
                                                  unsigned int n_substeps, const char* format,
                                                  unsigned int n_subsamples, unsigned int subsample_depth,
                                                  bool subsample_with_replacement, unsigned int pcoa_dims,
                                                  unsigned int permanova_perms, const char* grouping_filename,
                                                  const char* grouping_columns, const char* mmap_dir);  /* api.hpp:592 */

/* Partial-result files (stripes [start, stop) of the full matrix, LZ4 blocks,
 * magic PARTIAL_MAGIC_V2) and their merge into a full matrix. */
SSU_EXTERN IOStatus write_partial(const char* filename, const partial_mat_t* result);          /* api.hpp:857 */
SSU_EXTERN IOStatus read_partial(const char* filename, partial_mat_t** result);                /* api.hpp:872 */
SSU_EXTERN IOStatus read_partial_header(const char* input_filename, partial_dyn_mat_t** result_out); /* api.hpp:887 */
SSU_EXTERN IOStatus read_partial_one_stripe(partial_dyn_mat_t* result, uint32_t stripe_idx);  /* api.hpp:903 */
SSU_EXTERN MergeStatus validate_partial(const partial_dyn_mat_t* const* partial_mats, int n_partials); /* api.hpp:908 */
SSU_EXTERN MergeStatus merge_partial_to_mmap_matrix(partial_dyn_mat_t** partial_mats, int n_partials,
                                                    const char* mmap_dir, mat_full_fp64_t** result);   /* api.hpp:975 */
SSU_EXTERN MergeStatus merge_partial_to_mmap_matrix_fp32(partial_dyn_mat_t** partial_mats, int n_partials,
                                                         const char* mmap_dir, mat_full_fp32_t** result); /* api.hpp:991 */
#ifdef __cplusplus
/* api.hpp:942, :958 declare these two without extern "C" (C++ linkage) */
MergeStatus merge_partial_to_matrix(partial_dyn_mat_t** partial_mats, int n_partials, mat_full_fp64_t** result);
MergeStatus merge_partial_to_matrix_fp32(partial_dyn_mat_t** partial_mats, int n_partials, mat_full_fp32_t** result);
#endif

/* ==================================================================
 * Extensions (no counterpart in the reference)
 * ================================================================== */

/* Faith PD on in-memory inputs, same semantics as faith_pd_one_off
 * (samples with total < 1 are dropped first). */
SSU_EXTERN ComputeStatus ssu_faith_pd_inmem(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                            r_vec** result);

/* Device-resident plan: inputs uploaded once to HBM, then executed any
 * number of times; the n x n result stays in device memory until copied.
 * For benchmarking and for callers that chain GPU work on the matrix.
 * plain_kernel = 0 selects the specialised kernels (exact int8 MFMA Gram for
 * unweighted, shared-node min-formulation for weighted), 1 the plain dense
 * pair-tile kernel for every method (a cross-check and a baseline). */
typedef struct ssu_plan ssu_plan_t;
SSU_EXTERN ComputeStatus ssu_plan_create_inmem(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                               const char* unifrac_method, bool variance_adjust, double alpha,
                                               bool bypass_tips, bool normalize_sample_counts,
                                               int plain_kernel, ssu_plan_t** plan);
SSU_EXTERN ComputeStatus ssu_plan_execute(ssu_plan_t* plan);
/* execute computing only the pairs of stripes [start, stop) (start = stop = 0:
 * every pair, as ssu_plan_execute); entries of other pairs are unspecified.
 * Later ssu_plan_execute calls compute the same selection. */
SSU_EXTERN ComputeStatus ssu_plan_execute_stripes(ssu_plan_t* plan, unsigned int start, unsigned int stop);
/* stripes [start, stop) of the result as (stop - start) rows of n doubles:
 * row s - start, element k = d(k, (k + s + 1) mod n) */
SSU_EXTERN ComputeStatus ssu_plan_copy_stripes(const ssu_plan_t* plan, unsigned int start, unsigned int stop,
                                               double* host_out);
/* element size of the result: 8 (fp64 method names, and variance-adjusted fp32
 * names on tables whose counts FP32 cannot hold: a value that is not an integer,
 * or two sample totals summing to 2^24 or more) or 4 */
SSU_EXTERN int ssu_plan_result_bytes(const ssu_plan_t* plan);
SSU_EXTERN ComputeStatus ssu_plan_copy_matrix(const ssu_plan_t* plan, void* host_out);
/* rows [row0, row0 + nrows) of the n x n result (nrows * n elements), for
 * callers that stream a large matrix out in bands; output_error when the band
 * does not lie inside the matrix */
SSU_EXTERN ComputeStatus ssu_plan_copy_rows(const ssu_plan_t* plan, unsigned int row0, unsigned int nrows,
                                            void* host_out);
/* last execute(): milliseconds per phase (embed, prep, main kernel, combine, total, posting-list share of
 * the FP32 weighted path) from HIP events */
SSU_EXTERN int ssu_plan_timings(const ssu_plan_t* plan, double* ms, int cap);
SSU_EXTERN uint64_t ssu_plan_pair_nodes(const ssu_plan_t* plan);   /* n(n-1)/2 * processed nodes */
/* (pair, node) terms the main kernel of the last execute() evaluated, padding
 * included; the shared-node weighted kernel counts only the visited nodes */
SSU_EXTERN uint64_t ssu_plan_kernel_terms(const ssu_plan_t* plan);
/* (pair, node) terms of the last execute() summed from posting lists (FP32
 * weighted path: nodes held by few samples), not counted in kernel_terms */
SSU_EXTERN uint64_t ssu_plan_sparse_terms(const ssu_plan_t* plan);
/* pairs the last execute() recomputed in FP64 because their FP32 distance was
 * small enough for the FP32 rounding of the proportions to reach 1e-6 relative
 * (FP32 weighted and generalized methods; 0 otherwise) */
SSU_EXTERN uint64_t ssu_plan_refined_pairs(const ssu_plan_t* plan);
/* the occupancy split the plan chose (SSU_GRAM_SPLIT / SSU_SPARSE_SPLIT, or the
 * cost model of ssu_choose_split): nodes held by fewer than this fraction of
 * the samples are summed from posting lists; 0: no split */
SSU_EXTERN double ssu_plan_split_theta(const ssu_plan_t* plan);
/* wall-clock milliseconds of the calling thread's last one_off_matrix_inmem*
 * call: inputs, tree (shear / collapse / postorder), plan (uploads, schedules,
 * occupancy splits), execute (device work and refinement; with the copy of the
 * result when the two overlap, copy then 0), copy (result to the host), total.
 * Returns the number of values written (at most cap, 6). */
SSU_EXTERN int ssu_last_call_timings(double* ms, int cap);
/* Device buffers freed by finished calls stay cached in the process (mapping
 * and unmapping multi-gigabyte buffers costs more than the kernels); this
 * returns every cached buffer to the runtime.  SSU_POOL_GIB caps the cache
 * (default 90% of device memory; 0: no cache). */
SSU_EXTERN void ssu_release_device_memory(void);
/* The occupancy-split threshold a plan picks (host only): kind 0 FP32 weighted,
 * 1 FP64 weighted, 2 FP32 generalized, 3 unweighted Gram, 4 FP64 generalized,
 * 5 FP32 VAW, 6 FP64 VAW (one-sided terms priced at n_samples distinct totals);
 * occupancy[k] = number of samples holding node k.  Returns the threshold (nodes held by fewer samples
 * are summed from posting lists; 0: no split) and the modelled milliseconds
 * with and without it. */
SSU_EXTERN unsigned int ssu_choose_split(int kind, unsigned int n_samples, const uint32_t* occupancy,
                                         unsigned int n_nodes, double* model_ms, double* model_ms_none);
SSU_EXTERN const char* ssu_plan_main_kernel(const ssu_plan_t* plan);
SSU_EXTERN void ssu_plan_destroy(ssu_plan_t** plan);

/* Gower-centred matrix, PCoA (randomized FSVD, k = dims + 2 probes, two power
 * levels) and PERMANOVA on an n x n matrix, on the device.  The reference runs
 * these in its third-party skbb library (src/skbio_alt.cpp:144-233,
 * src/skbio_alt.hpp:22-52: su::mat_to_centered, su::pcoa, su::permanova),
 * which exports no C entry point of its own.  samples is n x dims, sample
 * major; dimensions past n come back zero.  grouping[i] in 0..groups-1; F and
 * p are NaN when there is one group or every sample is its own group.  Random
 * numbers follow the ssu_set_random_seed stream. */
SSU_EXTERN ComputeStatus ssu_center_matrix_fp64(const double* mat, unsigned int n, double* centered);
SSU_EXTERN ComputeStatus ssu_center_matrix_fp32(const float* mat, unsigned int n, float* centered);
/* Allocation-free Gower centring: no counterpart in the reference.  For
 * callers that cannot allocate device memory (work under HIP graph capture,
 * tight serving loops).  mat and centered are FP64 n x n row-major matrices in
 * device memory (they may be the same buffer); the work is enqueued on stream
 * (NULL: the null stream) and the call returns without waiting for it.  The
 * row statistics live in a fixed device array of the library (512 KiB per
 * GPU, held until the process ends), so n is at most 65,536 (larger n:
 * output_error, nothing enqueued), and calls that overlap on one GPU race.
 * The first call on a device looks the array up (hipGetSymbolAddress): make
 * it outside a capture.  The result equals ssu_center_matrix_fp64's. */
SSU_EXTERN ComputeStatus ssu_center_matrix_fp64_noalloc(const double* mat, unsigned int n, double* centered,
                                                        void* stream);
SSU_EXTERN ComputeStatus ssu_pcoa_fp64(const double* mat, unsigned int n, unsigned int dims, double* eigenvalues,
                                       double* samples, double* proportion_explained);
SSU_EXTERN ComputeStatus ssu_pcoa_fp32(const float* mat, unsigned int n, unsigned int dims, float* eigenvalues,
                                       float* samples, float* proportion_explained);
SSU_EXTERN ComputeStatus ssu_permanova_fp64(const double* mat, unsigned int n, const uint32_t* grouping,
                                            unsigned int n_perm, double* fstat, double* pvalue);
SSU_EXTERN ComputeStatus ssu_permanova_fp32(const float* mat, unsigned int n, const uint32_t* grouping,
                                            unsigned int n_perm, float* fstat, float* pvalue);

/* Rarefaction replicates on the device (the reference subsamples on the host,
 * src/biom_subsampled.cpp:155-321, one table per call).  Samples with a total
 * below depth are dropped; each of n_reps replicates draws every kept sample
 * down to depth (with replacement: multinomial in proportion to the counts;
 * without: depth of the sample's whole-count items), then computes every listed
 * method's full matrix.  The first execute runs eagerly and captures the whole
 * sequence into a HIP graph, later executes replay it.  Random numbers come from
 * a counter-based generator seeded with `seed`: statistical, not bitwise,
 * parity with the reference's sampler.  bypass_tips is not supported here. */
typedef struct ssu_rarefy ssu_rarefy_t;
SSU_EXTERN ComputeStatus ssu_rarefy_create_inmem(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                                 const char* const* unifrac_methods, int n_methods,
                                                 bool variance_adjust, double alpha, bool normalize_sample_counts,
                                                 unsigned int depth, bool with_replacement, uint64_t seed,
                                                 unsigned int n_reps, ssu_rarefy_t** out);
SSU_EXTERN ComputeStatus ssu_rarefy_execute(ssu_rarefy_t* batch);
/* kept samples / observations / CSR entries of the rarefied tables */
SSU_EXTERN unsigned int ssu_rarefy_n_samples(const ssu_rarefy_t* batch);
SSU_EXTERN unsigned int ssu_rarefy_n_obs(const ssu_rarefy_t* batch);
SSU_EXTERN uint64_t ssu_rarefy_nnz(const ssu_rarefy_t* batch);
/* original sample index of each kept sample, original observation row of each
 * kept row, and the kept CSR structure (indptr[n_obs+1], indices[nnz]) */
SSU_EXTERN void ssu_rarefy_kept(const ssu_rarefy_t* batch, uint32_t* sample_index, uint32_t* obs_index,
                                uint32_t* indptr, uint32_t* indices);
/* replicate rep's counts, aligned with the kept CSR (nnz values) */
SSU_EXTERN ComputeStatus ssu_rarefy_copy_counts(const ssu_rarefy_t* batch, unsigned int rep, double* data);
/* replicate rep's matrix of method `method` (index into unifrac_methods), n x n
 * doubles for "_fp64" names, floats otherwise */
SSU_EXTERN ComputeStatus ssu_rarefy_copy_matrix(const ssu_rarefy_t* batch, int method, unsigned int rep, void* host_out);
SSU_EXTERN double ssu_rarefy_last_ms(const ssu_rarefy_t* batch);     /* device time of the last execute */
SSU_EXTERN int ssu_rarefy_graph_used(const ssu_rarefy_t* batch);     /* 1: the last execute replayed the graph */
SSU_EXTERN uint64_t ssu_rarefy_pair_nodes(const ssu_rarefy_t* batch);
/* pairs of the last execute recomputed in FP64 (near-duplicate refinement of
 * the FP32 weighted and generalized methods), over every replicate */
SSU_EXTERN uint64_t ssu_rarefy_refined_pairs(const ssu_rarefy_t* batch);
SSU_EXTERN void ssu_rarefy_destroy(ssu_rarefy_t** batch);
/* Several rarefaction batches (tables of uneven sample counts) in one HIP graph
 * launch.  Extension, no counterpart in the reference, whose tables run one call
 * at a time.  Each batch's captured graph is a child node of one parent graph,
 * with no dependence between them, so the tables run side by side on the device.
 * Batches not yet executed run their first, eager execute inside create.  The
 * batches stay owned by the caller and must outlive the group.  After a group
 * execute, each batch's matrices, counts and refined-pair count read as after
 * its own execute; its last_ms reads 0, and the group's last_ms times the
 * whole launch. */
typedef struct ssu_rarefy_group ssu_rarefy_group_t;
SSU_EXTERN ComputeStatus ssu_rarefy_group_create(ssu_rarefy_t* const* batches, int n_batches,
                                                 ssu_rarefy_group_t** out);
SSU_EXTERN ComputeStatus ssu_rarefy_group_execute(ssu_rarefy_group_t* group);
SSU_EXTERN double ssu_rarefy_group_last_ms(const ssu_rarefy_group_t* group);
SSU_EXTERN void ssu_rarefy_group_destroy(ssu_rarefy_group_t** group);

/* Host-side tree handling only (no device): sheared + collapsed postorder
 * arrays of tree_data restricted to obs_ids.  Arrays hold n_parens/2 entries.
 * Returns the number of non-root nodes, or -1. */
SSU_EXTERN int ssu_postorder(const support_bptree_t* tree_data, int n_obs, const char* const* obs_ids,
                             int32_t* parent, double* length, int32_t* is_leaf, int32_t* obs_row);
/* Newick parse only: fills structure / lengths (n_parens entries) and copies
 * names into names_buf[i] (name_cap bytes each).  Returns n_parens or -n_parens if cap is too small. */
SSU_EXTERN int ssu_newick_to_bp(const char* newick, int cap, uint8_t* structure, double* lengths, char** names_buf,
                                int name_cap);
/* 1 if a HIP device is visible */
/* number of samples of a BIOM file (its /sample/ids), or -1; for the ssu
 * CLI's partial-report mode */
SSU_EXTERN long ssu_biom_n_samples(const char* biom_filename);
SSU_EXTERN int ssu_device_available(void);
SSU_EXTERN const char* ssu_last_error(void);

#endif /* SSU_AMD_H */
