// mf_io.h -- plotfile and checkpoint files of several levels (mf_io.hip), for
// the two-level driver.
#pragma once
#include <vector>

#include "mesoflow_amd.h"

namespace mfio {

// one level's part of the files: its geometry, boxes, component count and
// state [nv][nz][ny][nx] over its whole domain (only the boxes' cells are used)
struct LevelData {
    mf_geom geom;
    std::vector<mf_box> boxes;
    int nv;
    const double *state;
};

// amrex::WriteMultiLevelPlotfile with refinement ratio `ratio` between levels
int write_plotfile_levels(const std::vector<LevelData> &levels, double time, const std::vector<int> &isteps,
                          int ratio, const char *dir, const char *const *varnames);
// WriteCheckpointFile: isteps, dt and t_new hold one entry per possible level
int write_checkpoint_levels(const std::vector<LevelData> &levels, const std::vector<double> &t_new,
                            const std::vector<double> &dt, const std::vector<int> &isteps, const char *dir,
                            const char *const *varnames);
// ReadCheckpointFile's Header (MF_ERR_UNSUPPORTED above max_finest)
int read_checkpoint_header(const char *dir, int max_finest, int *finest, std::vector<int> &isteps,
                           std::vector<double> &dt, std::vector<double> &t_new, std::vector<std::vector<mf_box>> &boxes);
// the state of level l (boxes as in the Header) into a dense array over geom's domain
int read_checkpoint_level(const char *dir, int l, const mf_geom &geom, const std::vector<mf_box> &boxes, int nv,
                          double *state, const char *const *varnames);

}  // namespace mfio
