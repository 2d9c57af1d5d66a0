// HpgemAPIDevice<DIM>: hpGEM's explicit linear solver API with the
// right-hand side, the mass solve and the Runge-Kutta loop on one MI355X
// through the hgdg C ABI (include/hgdg.h).
//
// A maintainer drops this header into hpGEM (it compiles against
// kernel/Base/HpgemAPILinear.h, HpgemAPISimplified.h, ConfigurationData.h of
// the reference tree as is) and links libhgdg.so. Applications keep
// subclassing an HpgemAPI class: they derive from HpgemAPIDevice instead of
// HpgemAPILinear, pick the operator by kind (upwind advection with a
// velocity, or the acoustic system with getCInv evaluated per element), and
// keep their own getInitialSolution / getExactSolution / getSourceTerm hooks,
// which still run on the host through the reference's integrators.
//
// What moves to the device (each override names the reference function it
// replaces): computeTimeDerivative, computeRightHandSide,
// solveMassMatrixEquations, scaleAndAddVector, scaleVector,
// computeOneTimeStep (both forms), computeErrorAndNormOfUpdate, and the time
// loop of solve, which downloads the state only at output frames. The
// initial projection keeps the reference's host integrals
// (integrateInitialSolution) and solves with the mass matrix on the device.
// Source terms are integrated on the host by the reference's
// addSourceTerm/addSourceTermAtBoundary and added through M^{-1} on the
// device, so k = M^{-1}(S w) + M^{-1} f exactly as computeTimeDerivative of
// HpgemAPILinear forms k = M^{-1}(S w + f).
//
// Conventions: hpGEM's ERROR logging on a failing call (the library itself
// never exits), vector ids as hpGEM's, the tableau owned by hpGEM
// (ptrButcherTableau_), state in ElementData's layout (iVB = iV nB + iB,
// basis functions in hpGEM's order). The device holds every time-integration
// vector; the elements' host vectors are refreshed by download().
#ifndef HGDG_HPGEM_API_DEVICE_H
#define HGDG_HPGEM_API_DEVICE_H

#include <cmath>
#include <cstdint>
#include <fstream>
#include <limits>
#include <string>
#include <tuple>
#include <map>
#include <vector>

#include "Base/Element.h"
#include "Base/HpgemAPILinear.h"
#include "Base/MeshManipulator.h"
#include "Base/Node.h"
#include "Geometry/PhysicalGeometryBase.h"
#include "Geometry/PointPhysical.h"
#include "Geometry/PointReference.h"
#include "Geometry/ReferenceGeometry.h"
#include "Logger.h"
#include "Output/TecplotDiscontinuousSolutionWriter.h"
#include "Output/VTKTimeDependentWriter.h"
#include "hgdg.h"

namespace hpgem {
namespace Base {

template <std::size_t DIM>
class HpgemAPIDevice : public HpgemAPILinear<DIM> {
   public:
    using typename HpgemAPIBase<DIM>::PointPhysicalT;
    using typename HpgemAPIBase<DIM>::PointReferenceT;

    // kind: HGDG_OP_ADVECTION_UPWIND (one unknown, velocity a) or
    // HGDG_OP_ACOUSTIC_CENTRAL (DIM + 1 unknowns v, s_0 .. s_{DIM-1}).
    HpgemAPIDevice(int kind, const std::size_t polynomialOrder,
                   const TimeIntegration::ButcherTableau* const ptrButcherTableau,
                   const LinearAlgebra::SmallVector<DIM>& velocity = LinearAlgebra::SmallVector<DIM>(),
                   int precision = HGDG_PREC_FP64, int device = 0, const bool useSourceTerm = false,
                   const bool useSourceTermAtBoundary = false)
        : HpgemAPILinear<DIM>(kind == HGDG_OP_ACOUSTIC_CENTRAL ? DIM + 1 : 1, polynomialOrder, ptrButcherTableau, 0,
                              useSourceTerm, useSourceTermAtBoundary),
          kind_(kind),
          precision_(precision),
          velocity_(velocity) {
        check(hgdg_ctx_create(&ctx_, device));
    }

    // Any linear application of HpgemAPILinear: its own integrand hooks
    // (computeIntegrandMassMatrix, computeIntegrandStiffnessMatrixAtElement /
    // AtFace) build the element and face matrices on the host once, through
    // the reference's createMassMatrices / createStiffnessMatrices, and the
    // device runs the time loop on them (hgdg_operator_set_matrices).
    struct MatrixOperator {};
    HpgemAPIDevice(MatrixOperator, const std::size_t numberOfUnknowns, const std::size_t polynomialOrder,
                   const TimeIntegration::ButcherTableau* const ptrButcherTableau, int precision = HGDG_PREC_FP64,
                   int device = 0, const bool useSourceTerm = false, const bool useSourceTermAtBoundary = false)
        : HpgemAPILinear<DIM>(numberOfUnknowns, polynomialOrder, ptrButcherTableau, 0, useSourceTerm,
                              useSourceTermAtBoundary),
          kind_(HGDG_OP_MATRIX),
          precision_(precision) {
        check(hgdg_ctx_create(&ctx_, device));
    }

    HpgemAPIDevice(const HpgemAPIDevice& other) = delete;
    HpgemAPIDevice& operator=(const HpgemAPIDevice& other) = delete;

    ~HpgemAPIDevice() override {
        if (ctx_ != nullptr) hgdg_ctx_destroy(ctx_);
    }

    // The acoustic system's 1/c^2 weight (AcousticWaveLinear::getCInv),
    // taken at each element's centre: the device holds one value per element.
    virtual double getCInv(const PointPhysicalT& /*pPhys*/) { return 1.0; }

    // Mesh, operator and device vectors. Replaces HpgemAPILinear::
    // tasksBeforeSolving (kernel/Base/HpgemAPILinear_Impl.h:500-506), which
    // assembles the element and face matrices on the host: the device applies
    // the operator matrix-free, so nothing is assembled.
    void tasksBeforeSolving() override {
        if (!uploaded_) upload();
    }

    // HpgemAPISimplified::setInitialSolution (HpgemAPISimplified_Impl.h:751-758):
    // the reference's own host integrals of the initial solution against the
    // basis, then the mass solve on the device.
    void setInitialSolution(const std::size_t solutionVectorId, const double startTime,
                            const std::size_t orderTimeDerivative) override {
        if (!uploaded_) upload();
        this->integrateInitialSolution(solutionVectorId, startTime, orderTimeDerivative);
        uploadVector(solutionVectorId, solutionVectorId);
        check(hgdg_solve_mass(ctx_, int(solutionVectorId)));
        download(solutionVectorId);
    }

    // HpgemAPILinear::solveMassMatrixEquations (HpgemAPILinear_Impl.h:143-156).
    void solveMassMatrixEquations(const std::size_t timeIntegrationVectorId) override {
        check(hgdg_solve_mass(ctx_, int(timeIntegrationVectorId)));
    }

    // HpgemAPILinear::computeRightHandSide (HpgemAPILinear_Impl.h:485-498):
    // r = S w (+ f).
    void computeRightHandSide(const std::vector<std::size_t> inputVectorIds,
                              const std::vector<double> coefficientsInputVectors, const std::size_t resultVectorId,
                              const double time) override {
        std::vector<int> in(inputVectorIds.begin(), inputVectorIds.end());
        check(hgdg_right_hand_side(ctx_, in.data(), coefficientsInputVectors.data(), int(in.size()),
                                   int(resultVectorId), time));
        if (hasSource()) {
            sourceVector(resultVectorId, time);
            check(hgdg_scale_and_add(ctx_, int(resultVectorId), scratchId_, 1.0));
        }
    }

    void computeRightHandSide(const std::size_t inputVectorId, const std::size_t resultVectorId,
                              const double time) override {
        computeRightHandSide(std::vector<std::size_t>{inputVectorId}, std::vector<double>{1.0}, resultVectorId, time);
    }

    // HpgemAPISimplified::computeTimeDerivative (HpgemAPISimplified_Impl.h:
    // 777-786): stage combination, operator and inverse mass in one launch.
    void computeTimeDerivative(const std::vector<std::size_t> inputVectorIds,
                               const std::vector<double> coefficientsInputVectors, const std::size_t resultVectorId,
                               const double time) override {
        std::vector<int> in(inputVectorIds.begin(), inputVectorIds.end());
        check(hgdg_time_derivative(ctx_, in.data(), coefficientsInputVectors.data(), int(in.size()),
                                   int(resultVectorId), time));
        if (hasSource()) {
            sourceVector(resultVectorId, time);
            check(hgdg_solve_mass(ctx_, scratchId_));
            check(hgdg_scale_and_add(ctx_, int(resultVectorId), scratchId_, 1.0));
        }
    }

    void computeTimeDerivative(const std::size_t inputVectorId, const std::size_t resultVectorId,
                               const double time) override {
        computeTimeDerivative(std::vector<std::size_t>{inputVectorId}, std::vector<double>{1.0}, resultVectorId, time);
    }

    // HpgemAPISimplified::scaleAndAddVector / scaleVector (:697-717).
    void scaleAndAddVector(const std::size_t vectorToChangeId, const std::size_t vectorToAddId,
                           const double scale) override {
        check(hgdg_scale_and_add(ctx_, int(vectorToChangeId), int(vectorToAddId), scale));
    }

    void scaleVector(const std::size_t timeIntegrationVectorId, const double scale) override {
        // v + (scale - 1) v, one fused multiply-add per coefficient
        check(hgdg_scale_and_add(ctx_, int(timeIntegrationVectorId), int(timeIntegrationVectorId), scale - 1.0));
    }

    // HpgemAPISimplified::computeOneTimeStep(time, dt) (:788-820); with a source
    // term the reference's stage loop runs over the device calls above.
    void computeOneTimeStep(double& time, const double dt) override {
        if (hasSource()) {
            HpgemAPILinear<DIM>::computeOneTimeStep(time, dt);
            return;
        }
        std::vector<double> a, b, e;
        tableau(a, b, e);
        std::vector<int> aux(this->auxiliaryVectorIds_.begin(), this->auxiliaryVectorIds_.end());
        check(hgdg_rk_steps(ctx_, int(this->solutionVectorId_), aux.data(), int(b.size()), a.data(), b.data(), dt, 1));
        time += dt;
    }

    // HpgemAPISimplified::computeOneTimeStep(time, maxRelativeError, dtMax)
    // (:822-861). The step-size estimate is kept per object (the reference
    // keeps one function-static value for the whole process).
    void computeOneTimeStep(double& time, const double maxRelativeError, const double dtMax) override {
        if (hasSource()) {
            HpgemAPILinear<DIM>::computeOneTimeStep(time, maxRelativeError, dtMax);
            return;
        }
        std::vector<double> a, b, e;
        tableau(a, b, e);
        std::vector<int> aux(this->auxiliaryVectorIds_.begin(), this->auxiliaryVectorIds_.end());
        const double dtInitial = ::hpgem::Base::dt.isUsed() ? ::hpgem::Base::dt.getValue() : dtMax;
        check(hgdg_rk_adaptive_step(ctx_, int(this->solutionVectorId_), aux.data(), int(b.size()), a.data(), b.data(),
                                    e.data(), int(this->ptrButcherTableau_->getOrder()), maxRelativeError, dtMax,
                                    dtInitial, &time, nullptr, nullptr, nullptr));
    }

    // HpgemAPISimplified::computeErrorAndNormOfUpdate (:718-745).
    std::tuple<double, double> computeErrorAndNormOfUpdate(double dt) override {
        std::vector<double> a, b, e;
        tableau(a, b, e);
        std::vector<int> aux(this->auxiliaryVectorIds_.begin(), this->auxiliaryVectorIds_.end());
        double err = 0, norm = 0;
        check(hgdg_step_error(ctx_, int(this->solutionVectorId_), aux.data(), int(b.size()), e.data(), dt, &err,
                              &norm));
        return std::make_tuple(err, norm);
    }

    void tasksAfterSolving() override { download(this->solutionVectorId_); }

    // HpgemAPISimplified::solve (:936-1074): the same step count rule (ceil(T /
    // dt), rounded up to a multiple of the frame count), output at every frame
    // (Tecplot and VTK, from the downloaded state), the adaptive loop when the
    // error option is used, the errors at the end. Between frames the steps
    // run back to back on the device (one graph replay each).
    bool solve(const double startTime, const double endTime, double dt, const std::size_t numberOfOutputFrames,
               bool doComputeError) override {
        this->checkBeforeSolving();
        this->registerVTKWriteFunctions();
        Output::VTKTimeDependentWriter<DIM> vtk(this->outputFileName_, this->meshes_[0], this->polynomialOrder_);
        std::string dims, vars = this->variableNames_.empty() ? std::string("u") : this->variableNames_[0];
        for (std::size_t d = 0; d < DIM; ++d) dims += std::to_string(d);
        for (std::size_t v = 1; v < this->variableNames_.size(); ++v) vars += "," + this->variableNames_[v];
        std::ofstream tecFile(this->outputFileName_ + ".dat");
        Output::TecplotDiscontinuousSolutionWriter<DIM> tecplot(tecFile, this->internalFileTitle_, dims, vars);

        const double T = endTime - startTime;
        const std::size_t frames = numberOfOutputFrames > 0 ? numberOfOutputFrames : 1;
        std::size_t steps = std::size_t(std::ceil(T / dt));
        steps += (frames - steps % frames) % frames;
        dt = T / double(steps);
        const bool adaptive = ::hpgem::Base::error.isUsed();
        const double maxRelativeError = ::hpgem::Base::error.getValue();
        if (adaptive && !::hpgem::Base::dt.isUsed()) dt = std::numeric_limits<double>::infinity();

        double time = startTime;
        this->tasksBeforeSolving();
        this->setInitialSolution(this->solutionVectorId_, time, 0);
        auto writeFrame = [&]() {
            download(this->solutionVectorId_);
            tecplot.write(this->meshes_[0], this->solutionTitle_, false, this, time);
            this->VTKWrite(vtk, time, this->solutionVectorId_);
        };
        writeFrame();
        const double frameDt = T / double(frames);
        for (std::size_t f = 1; f <= frames; ++f) {
            const double frameEnd = f == frames ? endTime : startTime + double(f) * frameDt;
            if (adaptive) {
                while (time < frameEnd - 1e-14)
                    computeOneTimeStep(time, maxRelativeError, std::min(dt, frameEnd - time));
            } else if (hasSource()) {
                for (std::size_t s = 0; s < steps / frames; ++s) computeOneTimeStep(time, dt);
            } else {
                std::vector<double> a, b, e;
                tableau(a, b, e);
                std::vector<int> aux(this->auxiliaryVectorIds_.begin(), this->auxiliaryVectorIds_.end());
                const int n = int(steps / frames);
                check(hgdg_rk_steps(ctx_, int(this->solutionVectorId_), aux.data(), int(b.size()), a.data(), b.data(),
                                    dt, n));
                time += double(n) * dt;
            }
            writeFrame();
            this->showProgress(time, f * (steps / frames));
        }
        this->tasksAfterSolving();
        if (doComputeError) {
            const LinearAlgebra::MiddleSizeVector::type totalError = this->computeTotalError(this->solutionVectorId_, endTime);
            logger(INFO, "Total error: %.", totalError);
            const LinearAlgebra::MiddleSizeVector maxError = this->computeMaxError(this->solutionVectorId_, endTime);
            for (std::size_t v = 0; v < this->configData_->numberOfUnknowns_; ++v)
                logger(INFO, "Maximum error %: %", this->variableNames_[v], maxError(v));
        }
        return true;
    }

    // Device vector `vec` -> the elements' host vectors (ElementData layout).
    void download(const std::size_t vec) {
        std::vector<double> all(static_cast<std::size_t>(nDof_));
        check(hgdg_state_get(ctx_, int(vec), all.data()));
        std::size_t off = 0;
        for (Element* el : this->meshes_[0]->getElementsList()) {
            LinearAlgebra::MiddleSizeVector& v = el->getTimeIntegrationVector(vec);
            for (std::size_t i = 0; i < v.size(); ++i) v[i] = all[off + i];
            off += v.size();
        }
    }

    hgdg_ctx* context() { return ctx_; }

   protected:
    // The elements' host vector `hostVec` -> device vector `deviceVec`.
    void uploadVector(const std::size_t hostVec, int deviceVec) {
        std::vector<double> all;
        all.reserve(static_cast<std::size_t>(nDof_));
        for (Element* el : this->meshes_[0]->getElementsList()) {
            const LinearAlgebra::MiddleSizeVector& v = el->getTimeIntegrationVector(hostVec);
            for (std::size_t i = 0; i < v.size(); ++i) all.push_back(std::real(v[i]));
        }
        check(hgdg_state_set(ctx_, deviceVec, all.data()));
    }

   private:
    bool hasSource() const { return this->useSourceTerm_ || this->useSourceTermAtBoundary_; }

    // f(t) integrated by the reference's addSourceTerm / addSourceTermAtBoundary
    // into the host vectors of `hostVec` (zeroed first; the device holds the
    // authoritative copy of every vector), then uploaded to the scratch vector.
    void sourceVector(const std::size_t hostVec, const double time) {
        for (Element* el : this->meshes_[0]->getElementsList()) {
            LinearAlgebra::MiddleSizeVector& v = el->getTimeIntegrationVector(hostVec);
            for (std::size_t i = 0; i < v.size(); ++i) v[i] = 0.0;
        }
        if (this->useSourceTerm_) this->addSourceTerm(hostVec, time, 0);
        if (this->useSourceTermAtBoundary_) this->addSourceTermAtBoundary(hostVec, time, 0);
        uploadVector(hostVec, scratchId_);
    }

    void tableau(std::vector<double>& a, std::vector<double>& b, std::vector<double>& e) const {
        const TimeIntegration::ButcherTableau* bt = this->ptrButcherTableau_;
        const std::size_t s = bt->getNumberOfStages();
        a.assign(s * s, 0.0);
        b.assign(s, 0.0);
        e.assign(s, 0.0);
        for (std::size_t i = 0; i < s; ++i) {
            b[i] = bt->getB(i);
            if (bt->hasErrorEstimate()) e[i] = bt->getErrorCoefficient(i);
            for (std::size_t j = 0; j < i; ++j) a[i * s + j] = bt->getA(i, j);
        }
    }

    static int shapeOf(const Element* el) {
        switch (el->getReferenceGeometry()->getGeometryType()) {
            case Geometry::ReferenceGeometryType::SQUARE:
                return HGDG_SHAPE_SQUARE;
            case Geometry::ReferenceGeometryType::CUBE:
                return HGDG_SHAPE_CUBE;
            case Geometry::ReferenceGeometryType::TETRAHEDRON:
                return HGDG_SHAPE_TETRAHEDRON;
            case Geometry::ReferenceGeometryType::LINE:
                return HGDG_SHAPE_LINE;
            case Geometry::ReferenceGeometryType::TRIANGLE:
                return HGDG_SHAPE_TRIANGLE;
            case Geometry::ReferenceGeometryType::TRIANGULARPRISM:
                return HGDG_SHAPE_PRISM;
            case Geometry::ReferenceGeometryType::PYRAMID:
                return HGDG_SHAPE_PYRAMID;
            default:
                return 0;
        }
    }

    // Polynomial order of an element from its basis size (hpGEM's H1 DG sets:
    // squares (p+1)^2, cubes (p+1)^3, tetrahedra (p+1)(p+2)(p+3)/6).
    static int32_t orderOf(int shape, std::size_t nb) {
        for (int32_t p = 1; p <= 8; ++p)
            if (std::size_t(hgdg_basis_size(shape, p)) == nb) return p;
        logger(ERROR, "hgdg: no polynomial order has % basis functions on this shape", nb);
        return 0;
    }

    // hpGEM's own mesh as hgdg_mesh_desc takes it: per element the vertex
    // coordinates in the reference geometry's node order and the topological
    // node ids (periodic images share a node).
    void upload() {
        if (kind_ == HGDG_OP_MATRIX) {
            uploadMatrices();
            return;
        }
        std::vector<double> X;
        std::vector<int64_t> nodes;
        std::vector<int32_t> order;
        std::vector<double> cinv;
        int shape = 0;
        for (Element* el : this->meshes_[0]->getElementsList()) {
            shape = shapeOf(el);
            const Geometry::PhysicalGeometryBase* geo = el->getPhysicalGeometry();
            for (std::size_t v = 0; v < el->getNumberOfNodes(); ++v) {
                const Geometry::PointPhysical<DIM>& x = geo->getLocalNodeCoordinates(v).template castDimension<DIM>();
                for (std::size_t d = 0; d < DIM; ++d) X.push_back(x[d]);
                nodes.push_back(int64_t(el->getNode(v)->getID()));
            }
            order.push_back(orderOf(shape, el->getNumberOfBasisFunctions()));
            if (kind_ == HGDG_OP_ACOUSTIC_CENTRAL) {
                const PointReferenceT& centre = el->getReferenceGeometry()->getCenter().template castDimension<DIM>();
                cinv.push_back(getCInv(el->referenceToPhysical(centre)));
            }
        }
        hgdg_mesh_desc m;
        m.shape = shape;
        m.dim = int32_t(DIM);
        m.n_elements = int64_t(order.size());
        m.coords = X.data();
        m.node_ids = nodes.data();
        m.order = order.data();
        m.uniform_order = 0;
        check(hgdg_mesh_upload(ctx_, &m));
        hgdg_op_desc op;
        op.kind = kind_;
        op.precision = precision_;
        op.velocity[0] = op.velocity[1] = op.velocity[2] = 0.0;
        for (std::size_t d = 0; d < DIM; ++d) op.velocity[d] = velocity_[d];
        op.c_inv = cinv.empty() ? nullptr : cinv.data();
        check(hgdg_operator_set(ctx_, &op));
        check(hgdg_mesh_info(ctx_, nullptr, nullptr, nullptr, &nDof_));
        // every time-integration vector hpGEM keeps on the elements, plus one
        // scratch vector for source terms
        const std::size_t nVectors = this->meshes_[0]->getElementsList().size() > 0
                                         ? (*this->meshes_[0]->getElementsList().begin())->getNumberOfTimeIntegrationVectors()
                                         : 0;
        scratchId_ = int(nVectors);
        check(hgdg_state_alloc(ctx_, int(nVectors) + 1));
        uploaded_ = true;
    }

    // HpgemAPILinear::tasksBeforeSolving (HpgemAPILinear_Impl.h:500-506) on
    // the host, with the application's integrands, then every matrix to the
    // device in hgdg_matrix_desc's layout (column-major, rows = test).
    void uploadMatrices() {
        HpgemAPILinear<DIM>::tasksBeforeSolving();
        std::map<const Element*, int64_t> index;
        std::vector<int32_t> size;
        std::vector<double> mass, stiffness;
        auto append = [](std::vector<double>& to, const LinearAlgebra::MiddleSizeMatrix& A) {
            for (std::size_t c = 0; c < A.getNumberOfColumns(); ++c)
                for (std::size_t r = 0; r < A.getNumberOfRows(); ++r) to.push_back(A(r, c));
        };
        for (Element* el : this->meshes_[0]->getElementsList()) {
            index[el] = int64_t(size.size());
            const LinearAlgebra::MiddleSizeMatrix& M = el->getElementMatrix(this->massMatrixID_);
            size.push_back(int32_t(M.getNumberOfRows()));
            append(mass, M);
            append(stiffness, el->getElementMatrix(this->stiffnessElementMatrixID_));
        }
        std::vector<int64_t> faceElements;
        std::vector<double> blocks;
        for (Face* face : this->meshes_[0]->getFacesList()) {
            const FaceMatrix& F = face->getFaceMatrix(this->stiffnessFaceMatrixID_);
            faceElements.push_back(index.at(face->getPtrElementLeft()));
            append(blocks, F.getElementMatrix(Side::LEFT, Side::LEFT));
            if (face->isInternal()) {
                faceElements.push_back(index.at(face->getPtrElementRight()));
                append(blocks, F.getElementMatrix(Side::LEFT, Side::RIGHT));
                append(blocks, F.getElementMatrix(Side::RIGHT, Side::LEFT));
                append(blocks, F.getElementMatrix(Side::RIGHT, Side::RIGHT));
            } else {
                faceElements.push_back(-1);
            }
        }
        hgdg_matrix_desc d;
        d.n_elements = int64_t(size.size());
        d.n_fields = int32_t(this->configData_->numberOfUnknowns_);
        d.block_size = size.data();
        d.mass = mass.data();
        d.stiffness = stiffness.data();
        d.n_faces = int64_t(faceElements.size() / 2);
        d.face_elements = faceElements.data();
        d.face_blocks = blocks.data();
        check(hgdg_operator_set_matrices(ctx_, &d, precision_));
        check(hgdg_mesh_info(ctx_, nullptr, nullptr, nullptr, &nDof_));
        const std::size_t nVectors = this->meshes_[0]->getElementsList().size() > 0
                                         ? (*this->meshes_[0]->getElementsList().begin())->getNumberOfTimeIntegrationVectors()
                                         : 0;
        scratchId_ = int(nVectors);
        check(hgdg_state_alloc(ctx_, int(nVectors) + 1));
        uploaded_ = true;
    }

    // hpGEM's error convention: a failing call is a logged error.
    static void check(int status) {
        if (status != HGDG_OK) logger(ERROR, "hgdg: %", hgdg_last_error());
    }

    hgdg_ctx* ctx_ = nullptr;
    int kind_;
    int precision_;
    LinearAlgebra::SmallVector<DIM> velocity_;
    int64_t nDof_ = 0;
    int scratchId_ = 0;
    bool uploaded_ = false;
};

}  // namespace Base
}  // namespace hpgem

#endif  // HGDG_HPGEM_API_DEVICE_H
