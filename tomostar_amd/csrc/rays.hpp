// The rows of a batch of rays (RayKernels.RayRow, Tomography/RayKernels.cs:15-27): per arrival its
// node list (ascending) and float32 lengths, resident in HBM as CSR in arrival order, plus TCalc and
// the hypocentre gradient on the host.
#pragma once

#include <vector>

#include "common.hpp"

struct ts_rays {
    ts_context *ctx = nullptr;
    int nrays = 0;
    ts_grid inversion{};              // the grid the rows' node indices refer to
    std::vector<int32_t> counts;      // nodes per row, -1: no row (stalled or unreachable)
    std::vector<int64_t> row_start;   // CSR over the arrivals that have rows (empty rows for the others)
    ts::dbuf<int32_t> nodes;              // device CSR of the rows
    ts::dbuf<float> lengths;
    std::vector<double> tcalc_grad;   // 4 per arrival: TCalc, Ge, Gn, Gz
    std::vector<int32_t> path_points; // points of each path (diagnostic)
    std::vector<double> outside;      // per arrival: time of the pieces outside the inversion grid (0: none)
    bool extended = false;            // traced with a background (pieces outside the grid timed, not attributed)
    double seconds = 0;
    std::unique_ptr<ts::HostStage> stage;  // pinned staging of large row reads (copy_to_host)
};

