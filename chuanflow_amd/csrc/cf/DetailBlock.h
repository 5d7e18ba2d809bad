// DetailBlock.h -- one 3D detail region of the 1D-3D coupling, evaluated on
// the GPU: the reference's DetailBlock (CoupledNetwork.h; coupledNetwork.cpp
// :644-926, :1080-1200) for the laminar incompressible case.
//
// The inlet and outlet are the patches typed "inlet" and "outlet"; the flow
// axis and direction come from the inlet's area vector.  lossAt(q) imposes
// the rate q on the inlet (a plug q/A, or a parabolic / 1/7-power profile
// over the wall-bounded transverse axes, normalised so that its area
// integral is the section area), runs SIMPLE on the device, warm-started
// from the block's previous fields, and measures the drop of the
// volume-weighted section-mean pressure between the first and the last
// cross-section, in Pa (rho times the kinematic p).
#pragma once

#include "PolyMesh.h"

#include "../../../include/cfx.h"

#include <string>
#include <vector>

namespace cf
{

struct BlockMeasure
{
    label cells = 0;
    int sweeps = 0, converged = 0;
    double imposedQ = 0, inletArea = 0, plugSpeed = 0;
    double dpPa = 0, fluxIn = 0, fluxOut = 0;
};

class DetailBlock
{
public:
    DetailBlock(cfx_ctx* ctx, const polyMesh& mesh, double nu, double rho, const std::string& inletShape);
    ~DetailBlock();
    DetailBlock(const DetailBlock&) = delete;
    DetailBlock& operator=(const DetailBlock&) = delete;

    // drop across the region at rate q (m3/s), SIMPLE with the given controls
    double lossAt(double q, const cfx_simple_controls& ctl);
    const BlockMeasure& lastMeasure() const { return measure_; }

    int axis() const { return axis_; }
    int flowSign() const { return flowSign_; }
    double inletArea() const { return inletArea_; }
    cfx_ctx* context() const { return ctx_; }
    std::size_t nSections() const { return sections_.size(); }
    const std::vector<label>& inletFaces() const { return inletFaces_; }
    const std::vector<double>& inletWeights() const { return inletWeights_; }
    cfx_mesh* deviceMesh() const { return dm_; }
    // host copies of the fields: U [3][n] (SoA), p [n], phi [nFaces]
    void fields(double* U, double* p, double* phi) const;

private:
    cfx_ctx* ctx_;
    polyMesh mesh_;
    double nu_, rho_;
    std::string shape_;
    label inletPatch_ = -1, outletPatch_ = -1;
    int axis_ = 0, flowSign_ = 1;
    double inletArea_ = 0;
    std::vector<std::vector<label>> sections_;
    std::vector<label> inletFaces_;
    std::vector<double> inletWeights_;
    cfx_mesh* dm_ = nullptr;
    double *dU_ = nullptr, *dp_ = nullptr, *dphi_ = nullptr, *dlc_ = nullptr, *dlk_ = nullptr;
    // measurement on the device: cell volumes; segments {first section, last section} of cells and
    // {inlet, outlet} of faces, as start offsets [3] and indices
    double* dV_ = nullptr;
    int64_t *dCellSeg_ = nullptr, *dCellIdx_ = nullptr, *dFaceSeg_ = nullptr, *dFaceIdx_ = nullptr;
    BlockMeasure measure_;

    Vec3 patchArea(label patch) const;
    void buildSections();
    void buildInletProfile();
    void applyBoundary(double q, std::vector<int32_t>& Ut, std::vector<double>& Uv, std::vector<int32_t>& pt,
                       std::vector<double>& pv);
    void measure(const cfx_simple_result& r, double q);
    void check(int rc) const;
};

} // namespace cf
