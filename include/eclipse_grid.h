/*
 * eclipse_grid.h -- ECLIPSE (GRDECL) grid and rock input for the MI355X
 * mimetic hybrid solver, host side, C-ABI.
 *
 * This replaces the ECLIPSE branch of the reference's set-up,
 *   setupGridAndProps (opm/porsol/common/setupGridAndProps.hpp:78-117):
 *   Dune::CpGrid::processEclipseFormat (dune-cornerpoint, not in the
 *   reference tree) for the grid, and ReservoirPropertyCommon::init(deck,
 *   global_cell, perm_threshold, ...) (opm/porsol/common/
 *   ReservoirPropertyCommon_impl.hpp:275-300) for the rock: the permeability
 *   keyword rules classifyPermeability / fillTensor / assignPermeability
 *   (:52-244, 682-736) and assignPorosity (:592-611).
 *
 * Supported decks: SPECGRID or DIMENS, COORD with vertical pillars on a
 * regular lattice, ZCORN with gap-free layers (faults between columns are
 * allowed: every overlap of two columns' layers on their shared vertical plane
 * becomes a face, the parts facing no cell become no-flow boundary faces with
 * id 7), all cells active, METRIC or FIELD units (FIELD lengths are converted
 * from feet).  Keywords outside the GRID and PROPS arrays are skipped;
 * reading stops at the first SOLUTION, SUMMARY, SCHEDULE or REGIONS section.
 * Other decks are rejected with an error.  dune-cornerpoint's own geometry
 * and face order are not reproduced.
 *
 * The grid comes out in the layout mhs_set_grid takes (cells i fastest, then
 * j, then k; per cell its half-faces: x-, x+, y-, y+ sides (sub-faces by
 * neighbour layer), top, bottom), with boundary ids 1..6 by side of the box
 * and 7 for faces exposed by a fault.  Lengths in metres, permeability in
 * m^2 (deck values in millidarcy).
 *
 * Every function returns 0 on success; on failure mhs_deck_last_error()
 * describes the cause.  A deck handle is used by one host thread at a time.
 */
#ifndef ECLIPSE_GRID_H
#define ECLIPSE_GRID_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhs_deck mhs_deck;

/* Reads a GRDECL / DATA file (Opm::Parser::parseFile in the reference's
 * set-up, setupGridAndProps.hpp:87-88).  *out is set even on failure when
 * non-null, so that mhs_deck_last_error can report the reason; destroy it. */
int mhs_deck_read_file(mhs_deck **out, const char *path);
/* The same from text in memory. */
int mhs_deck_read_text(mhs_deck **out, const char *text);
int mhs_deck_destroy(mhs_deck *d);
const char *mhs_deck_last_error(const mhs_deck *d);

/* Logical dimensions (SPECGRID or DIMENS). */
int mhs_deck_dims(mhs_deck *d, int dims[3]);

/* Builds the corner-point grid (CpGrid::processEclipseFormat) and reports its
 * size; the arrays follow with mhs_deck_grid_arrays. */
int mhs_deck_build_grid(mhs_deck *d, int *num_cells, int *num_half_faces);
/* The grid of mhs_deck_build_grid: facepos[nc+1], neighbour[nhf] (-1 on the
 * boundary), boundary_id[nhf] (0 inside), cell_volume[nc],
 * cell_centroid[3 nc], face_area[nhf], face_centroid[3 nhf],
 * face_normal[3 nhf] (unit outer normals), ijk[3 nc] (logical indices,
 * CpGrid::globalCell is i + nx (j + ny k)), corners[24 nc] (the eight
 * corners of each cell in VTK hexahedron order: the top ring (i, j),
 * (i+1, j), (i+1, j+1), (i, j+1), then the bottom ring; for output).  Any
 * pointer may be NULL. */
int mhs_deck_grid_arrays(mhs_deck *d, int *facepos, int *neighbour, int *boundary_id,
                         double *cell_volume, double *cell_centroid, double *face_area,
                         double *face_centroid, double *face_normal, int *ijk, double *corners);

/* ReservoirPropertyCommon::init(deck, global_cell, perm_threshold, ...)
 * (ReservoirPropertyCommon_impl.hpp:275-300) for num_cells cells with
 * global_cell[num_cells] (NULL: the identity): perm[9 nc] row-major tensors
 * in m^2 (zero without permeability keywords; diagonal raised to
 * perm_threshold), porosity[nc] (1 without PORO), *kind: 0 none, 1 scalar,
 * 2 diagonal, 3 tensor.  Either output may be NULL.  Errors carry the
 * reference's texts ("Invalid set of permeability fields given.", ...). */
int mhs_deck_rock(mhs_deck *d, int num_cells, const int *global_cell, double perm_threshold,
                  double *perm, double *porosity, int *kind);

#ifdef __cplusplus
}
#endif

#endif /* ECLIPSE_GRID_H */
