// A client of the GooFit C++ API, written the way the reference's examples use it
// (examples/simpleFit/SimpleFit.cpp, examples/addition/addition.cpp,
// examples/dalitz/dalitz.cpp:281-335): Observables and Variables, the PDF classes,
// an UnbinnedDataSet filled with setValue + addEvent, setData / setDataSize,
// calculateNLL and a FitManager fit.  It is compiled with g++ against include/
// and linked to libgoofit.so only; the tests run it and compare its output with
// the golden fixtures.
//
// usage:
//   goofit_client add1d  <events.txt> [fit]
//       the 1-D AddPdf: 0.3 Gaussian(0.3, 0.8) + Exp(-0.4) on [-5, 5], 1000 bins;
//       events one value per line
//   goofit_client dalitz <model.txt> <events.txt> <grid> [fit]
//       an Amp3Body of D0 -> pi0 pi+ pi-; model.txt holds one resonance per line,
//       "name kind mass width spin pair re im" (kind RBW, GS or NONRES; the first
//       line's coupling is fixed, the others float when fitting); events "m12 m13"
//       per line; grid = bins of m12 and m13 on [0, 3]
// Output lines "nll <value>", and after a fit "valid <0|1>", "fval <value>",
// "param <name> <value> <error>".
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <goofit/Error.h>
#include <goofit/FitManager.h>
#include <goofit/PDFs/basic/ExpPdf.h>
#include <goofit/PDFs/basic/GaussianPdf.h>
#include <goofit/PDFs/basic/PolynomialPdf.h>
#include <goofit/PDFs/combine/AddPdf.h>
#include <goofit/PDFs/physics/Amp3Body.h>
#include <goofit/PDFs/physics/DalitzPlotHelpers.h>
#include <goofit/PDFs/physics/resonances/GS.h>
#include <goofit/PDFs/physics/resonances/NonRes.h>
#include <goofit/PDFs/physics/resonances/RBW.h>
#include <goofit/UnbinnedDataSet.h>
#include <goofit/Variable.h>

using namespace GooFit;

namespace {

std::vector<std::vector<double>> read_rows(const std::string &path, size_t ncols) {
    std::ifstream in(path);
    if(!in)
        throw GeneralError("cannot open " + path);
    std::vector<std::vector<double>> rows;
    std::string line;
    while(std::getline(in, line)) {
        std::istringstream ss(line);
        std::vector<double> r(ncols);
        bool ok = true;
        for(size_t k = 0; k < ncols && ok; k++)
            ok = static_cast<bool>(ss >> r[k]);
        if(ok)
            rows.push_back(r);
    }
    return rows;
}

void print_fit(FitManager &fm, const std::vector<Variable *> &vars) {
    auto min = fm.fit();
    std::printf("valid %d\nfval %.17g\n", min.IsValid() ? 1 : 0, min.Fval());
    for(const Variable *v : vars)
        std::printf("param %s %.17g %.17g\n", v->getName().c_str(), v->getValue(), v->getError());
}

int run_add1d(int argc, char **argv) {
    Observable xvar{"xvar", -5, 5};
    xvar.setNumBins(1000);
    Variable mean{"mean", 0.3, 0.01, -2, 2};
    Variable sigma{"sigma", 0.8, 0.01, 0.1, 3};
    Variable alpha{"alpha", -0.4, 0.01, -2, -0.01};
    Variable frac{"frac", 0.3, 0.01, 0, 1};

    GaussianPdf signal{"signal", xvar, mean, sigma};
    ExpPdf background{"background", xvar, alpha};
    AddPdf total{"total", frac, &signal, &background};

    UnbinnedDataSet data{xvar};
    for(const auto &r : read_rows(argv[2], 1)) {
        xvar.setValue(r[0]);
        data.addEvent();
    }
    total.setData(&data);
    std::printf("nll %.17g\n", total.calculateNLL());
    if(argc > 3 && std::string(argv[3]) == "fit") {
        FitManager fm{&total};
        fm.setVerbosity(0);
        print_fit(fm, {&mean, &sigma, &alpha, &frac});
    }
    return 0;
}

int run_dalitz(int argc, char **argv) {
    const int grid = std::stoi(argv[4]);
    Observable m12{"m12", 0, 3};
    Observable m13{"m13", 0, 3};
    EventNumber eventNumber{"eventNumber"};
    m12.setNumBins(grid);
    m13.setNumBins(grid);

    DecayInfo3 dtop0pp;
    dtop0pp.motherMass          = 1.8648;
    dtop0pp.daug1Mass           = 0.13498;
    dtop0pp.daug2Mass           = 0.13957;
    dtop0pp.daug3Mass           = 0.13957;
    dtop0pp.meson_radius        = 1.5;
    dtop0pp.mother_meson_radius = 5.0;  // set explicitly: the examples leave it uninitialised

    std::vector<Variable *> floating;
    std::ifstream model(argv[2]);
    if(!model)
        throw GeneralError(std::string("cannot open ") + argv[2]);
    std::string line;
    bool first = true;
    while(std::getline(model, line)) {
        std::istringstream ss(line);
        std::string name, kind;
        double mass, width, re, im;
        unsigned int spin, pair;
        if(line.find_first_not_of(" \t\r") == std::string::npos)
            continue;
        if(!(ss >> name >> kind >> mass >> width >> spin >> pair >> re >> im))
            throw GeneralError("malformed resonance line in " + std::string(argv[2]) + ": " + line);
        // the first coupling fixes the overall complex scale; the others float
        Variable *ar = first ? new Variable(name + "_amp_real", re) : new Variable(name + "_amp_real", re, 0.01, 0, 0);
        Variable *ai = first ? new Variable(name + "_amp_imag", im) : new Variable(name + "_amp_imag", im, 0.01, 0, 0);
        if(!first) {
            floating.push_back(ar);
            floating.push_back(ai);
        }
        first = false;
        ResonancePdf *r = nullptr;
        if(kind == "NONRES")
            r = new Resonances::NonRes(name, *ar, *ai);
        else if(kind == "GS")
            r = new Resonances::GS(name, *ar, *ai, Variable(name + "_mass", mass), Variable(name + "_width", width),
                                   spin, pair);
        else
            r = new Resonances::RBW(name, *ar, *ai, Variable(name + "_mass", mass), Variable(name + "_width", width),
                                    spin, pair);
        dtop0pp.resonances.push_back(r);
    }

    if(dtop0pp.resonances.empty())
        throw GeneralError(std::string("no resonance in ") + argv[2]);

    Variable constantOne{"constantOne", 1};
    Variable constantZero{"constantZero", 0};
    std::vector<Observable> observables{m12, m13};
    std::vector<Variable> offsets{constantZero, constantZero};
    std::vector<Variable> coefficients{constantOne};
    PolynomialPdf eff{"constantEff", observables, coefficients, offsets, 0};
    Amp3Body signal{"signalPDF", m12, m13, eventNumber, dtop0pp, &eff};

    UnbinnedDataSet data{{m12, m13, eventNumber}};
    for(const auto &r : read_rows(argv[3], 2)) {
        m12.setValue(r[0]);
        m13.setValue(r[1]);
        eventNumber.setValue(data.getNumEvents());
        data.addEvent();
    }
    signal.setData(&data);
    signal.setDataSize(data.getNumEvents());
    std::printf("nll %.17g\n", signal.calculateNLL());
    if(argc > 5 && std::string(argv[5]) == "fit") {
        FitManager fm{&signal};
        fm.setVerbosity(0);
        print_fit(fm, floating);
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if(mode == "add1d" && argc > 2)
            return run_add1d(argc, argv);
        if(mode == "dalitz" && argc > 4)
            return run_dalitz(argc, argv);
        std::fprintf(stderr, "usage: %s add1d <events> [fit] | dalitz <model> <events> <grid> [fit]\n", argv[0]);
        return 2;
    } catch(const GeneralError &e) {
        std::fprintf(stderr, "GeneralError: %s\n", e.what());
        return 1;
    }
}
