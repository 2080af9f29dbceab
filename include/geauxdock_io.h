/*
 * geauxdock_io.h -- GeauxDock's input files as a gd_complex_desc.
 *
 * Reads the four inputs of the reference's `dock` driver (src/dock.C:21-94,
 * src/dock.C:104-186) and applies its host-side derivations, so a complex
 * prepared for the reference runs unchanged on this engine:
 *
 *   ligand SDF      default conformer + ENSEMBLE_COORDS conformers whose
 *                   ENSEMBLE_RMSD exceeds 0.1, OB_ATOM_TYPES, OB_ATOMIC_CHARGES,
 *                   each conformer moved to its own centroid
 *                   (src/load.C:157-447);
 *   protein PDB     coarse points per residue (CA, peptide, one or two
 *                   side-chain centroids) averaged per MODEL
 *                   (src/load.C:931-1313), then CopyProteinResidue
 *                   (src/util.C:574-597);
 *   force field .ff CENTER (pocket centre), KDE points, PSP table and MCS
 *                   templates (src/load.C:1326-1475);
 *   parameters      `paras` (VDW, PMF, HDB, ELE, HPP, HPL, PLJ, PEL, KDE,
 *                   WEI, NOA, NOB; src/load.C:1477-1660), optionally a
 *                   weight file and normalization files (src/load.C:1662-1794),
 *                   then OptimizeEnepara (src/util.C:717-768);
 *   temperatures    SetTemperature's geometric ladder (src/util.C:797-837).
 *
 * Errors are reported through the status and a message buffer, never exit().
 */
#ifndef GEAUXDOCK_IO_H
#define GEAUXDOCK_IO_H

#include "geauxdock.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gd_input_files {
  const char *ligand_sdf;  /* -l */
  const char *protein_pdb; /* -p */
  const char *lhm_ff;      /* -s */
  const char *enepara;     /* --para */
  const char *complex_id;  /* --id: MCS records are matched against it */
  const char *weight;      /* optional: one weight per line (loadWeight) */
  const char *nor_a;       /* optional: normalization a, one per line */
  const char *nor_b;       /* optional: normalization b, one per line */
} gd_input_files;

typedef struct gd_dock_params {
  int n_tmp;          /* --nt */
  float floor_temp;   /* --floor_temp */
  float ceiling_temp; /* --ceiling_temp */
  float ts, rs;       /* -t, -r: translation and rotation move scales */
  int move_kind;      /* GD_MOVE_UNIFORM (the reference) or GD_MOVE_GAUSSIAN */
} gd_dock_params;

typedef struct gd_loaded gd_loaded;

/* Parse and derive. *out owns every array the descriptor points to. */
gd_status gd_load_complex(const gd_input_files *files, const gd_dock_params *params,
                          gd_loaded **out, char *err, size_t err_len);
/* The complex, ready for gd_add_complex (valid until gd_loaded_free). */
const gd_complex_desc *gd_loaded_desc(const gd_loaded *l);
/* Extra facts of the load: ligand records in the SDF, raw ENSEMBLE_TOTAL,
 * coarse residues, KDE points dropped for an unknown atom type. */
void gd_loaded_info(const gd_loaded *l, int *sdf_sections, int *ensemble_total, int *residues,
                    int *kde_dropped);
void gd_loaded_free(gd_loaded *l);

/* One SDF record as the file holds it, for checks and tools (the reference's
 * readLigandSections, writeLigProperty, writeLigAtomProperty,
 * writeDefaultLigAtomCoord, getLigEnsembleTotal/Coords/Rmsd, writeEnsAtomCoord
 * and moveLigand2ItsCenterFrame; src/load.C:157-447). Atom types are not
 * required here (gd_load_complex requires them). */
typedef struct gd_sdf_record {
  int records;        /* $$$$-terminated records in the file */
  int lna, lnb;       /* atoms and bonds of the chosen record */
  int typed;          /* 1 when the record has OB_ATOM_TYPES */
  int ensemble_total; /* the ENSEMBLE_TOTAL field (0 when absent) */
  int ensemble_lines; /* ENSEMBLE_COORDS lines, = ENSEMBLE_RMSD values */
  int conformers;     /* docked conformers: the atom block + lines with RMSD > 0.1 */
} gd_sdf_record;

/* record: 0-based. xyz (optional): conformers in file order, the atom block
 * first and then every ENSEMBLE_COORDS line, each as x[lna] y[lna] z[lna];
 * at most max_conf are written. centre_each: move each to its own centroid,
 * as gd_load_complex does. rmsd (optional): up to max_rmsd ENSEMBLE_RMSD
 * values. types (optional, lna): ligand type codes, -1 where unknown or
 * absent. charges (optional, lna): OB_ATOMIC_CHARGES. */
gd_status gd_sdf_read(const char *path, int record, int centre_each, gd_sdf_record *info,
                      float *xyz, int max_conf, float *rmsd, int max_rmsd, int *types,
                      float *charges, char *err, size_t err_len);

/* ---- trajectory output ------------------------------------------------------ */

/* printStates for accepted steps (src/util.C:1504-1550): header
 * "lig prt cms rmsd mv1..mv6 vdw ele pmf psp hdb hpc kde lhm edst ener", one
 * line per record, values with 4 decimals. */
gd_status gd_write_states_csv(const char *path, const gd_record *recs, int64_t n, char *err,
                              size_t err_len);
/* printStates for medoids (src/util.C:1464-1502): move vector, nine terms,
 * cms, rmsd, total and the cluster size, values with 6 decimals. */
gd_status gd_write_medoids_csv(const char *path, const gd_record *recs, const int *cluster_size,
                               int64_t n, char *err, size_t err_len);

/* ---- record file ---------------------------------------------------------------
 * Replaces the reference's per-dump HDF5 LigRecord files (DumpLigRecord /
 * ReadLigRecord, src/hdf5io.C:11-55; layout
 * src/hdf5_datastruct_ligrecord_create.h:53-84) with one flat file: a 64-byte
 * header ("GDREC001", record size 92, version 1, record count) followed by
 * gd_record entries in append order. */
gd_status gd_records_write(const char *path, int append, const gd_record *recs, int64_t n,
                           char *err, size_t err_len);
gd_status gd_records_count(const char *path, int64_t *n, char *err, size_t err_len);
gd_status gd_records_read(const char *path, int64_t first, int64_t n, gd_record *out, char *err,
                          size_t err_len);

/* ---- trace tool (src/trace.C:90-119) --------------------------------------------- */

/* PlaceLigand (src/util.C:456-497): coordinates of ligand conformer idx_lig
 * of a complex under move vector mm (translation from the pocket centre,
 * Euler angles); x, y, z hold lna values. */
gd_status gd_pose_coords(const gd_complex_desc *d, int idx_lig, const float mm[6], float *x,
                         float *y, float *z);
/* replaceLigandCoords (src/util.C:504-549): ligand_sdf with the first
 * record's atom coordinates replaced, written to out_path. */
gd_status gd_write_pose_sdf(const char *ligand_sdf, int lna, const float *x, const float *y,
                            const float *z, const char *out_path, char *err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif
