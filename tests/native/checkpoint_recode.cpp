// Test harness (host only, no device): reads a preset's SWST checkpoint
// bytes, then writes them again twice: verbatim, and with the suspension
// section restated from the decoded state (as a checkpoint taken after a
// step is written). Prints "verbatim <0|1> restated <0|1>" (1: the bytes
// equal the input). The definition is the preset's with its hinges in the
// build's colour order, the order the reference's Structure received.
//   checkpoint_recode <preset> <bytes-file> <substep-dt>
#include "checkpoint_codec.hpp"
#include "schedule.hpp"
#include "simwing_xpbd.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <vector>

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    sw_scene* scene = nullptr;
    if (sw_scene_generate(argv[1], &scene) != SW_OK) return 3;
    sw_assembly_settings settings{};
    settings.seam_coincidence_tolerance = 1.0e-9;
    settings.suspension_iterations = 48;
    settings.suspension_attachment_tolerance = 1.0e-10;
    settings.suspension_minimum_line_length = 1.0e-10;
    settings.suspension_maximum_line_residual = 0.10;
    settings.suspension_maximum_control_work = 1.0e6;
    sw_assembly* assembly = nullptr;
    if (sw_scene_assemble(scene, &settings, &assembly) != SW_OK) return 4;
    sw_definition view;
    if (sw_assembly_definition(assembly, &view) != SW_OK) return 5;
    swx::Definition d = swx::Definition::fromView(view);
    const swx::ColourOrder hinges = swx::colourDihedrals(d);
    swx::Definition ordered = d;
    for (std::size_t k = 0; k < hinges.order.size(); ++k) {
        const std::size_t o = hinges.order[k];
        for (int c = 0; c < 4; ++c) ordered.dihedralNodes[4 * k + c] = d.dihedralNodes[4 * o + c];
        ordered.dihedralRestAngle[k] = d.dihedralRestAngle[o];
        ordered.dihedralCompliance[k] = d.dihedralCompliance[o];
    }
    std::ifstream file(argv[2], std::ios::binary);
    const std::vector<std::uint8_t> bytes((std::istreambuf_iterator<char>(file)), std::istreambuf_iterator<char>());
    swx::CheckpointImage image;
    swx::DeviceWing::State& s = image.state;
    const std::size_t n = ordered.nodeCount(), m = ordered.membraneCount();
    s.position.resize(3 * n);
    s.previous.resize(3 * n);
    s.velocity.resize(3 * n);
    s.pressure.resize(ordered.triangleCount());
    s.constraintLambda.resize(ordered.constraintCount());
    s.membraneLambda.resize(3 * m);
    s.membraneResultant.resize(3 * m);
    s.membraneResidual.resize(m);
    s.dihedralLambda.resize(ordered.dihedralCount());
    s.payload.assign(26, 0.0);
    s.segmentLambda.assign(ordered.suspension ? ordered.suspension->segmentIds.size() : 0, 0.0);
    try {
        swx::decodeStructureCheckpoint(ordered, swx::definitionFingerprint(ordered), bytes.data(), bytes.size(), image);
        const bool verbatim = swx::encodeStructureCheckpoint(ordered, image) == bytes;
        image.suspensionSection.clear();
        image.substepDt = std::atof(argv[3]);
        const bool restated = swx::encodeStructureCheckpoint(ordered, image) == bytes;
        std::printf("verbatim %d restated %d\n", verbatim ? 1 : 0, restated ? 1 : 0);
    } catch (const std::exception& e) {
        std::printf("error %s\n", e.what());
        return 1;
    }
    return 0;
}
