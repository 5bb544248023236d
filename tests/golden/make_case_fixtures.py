"""Copy the data files of four reference tutorial cases into fixtures for the case I/O tests.

Sources (data only: meshes, field files and dictionaries of HopeFOAM-0.1/tutorials/DG/2D):
  * isentropicVortex: the Fluent meshes vortex0256.msh and vortex1024.msh, system/{controlDict,dgSolution,dgSchemes},
    0/{rho,rhoU,Ener} (uniform initial fields; Wall fixedValue, frontAndBackPlanes empty);
  * cylinder: the one-layer prism polyMesh constant/polyMesh/{points,faces,owner,neighbour,boundary},
    system/{controlDict,dgSolution}, 0/{U,p} (inlet/channel/cylinder fixedValue, outlet zeroGradient);
  * pearsonVortex (the projection solver dgChorinFoam): the Fluent meshes vortexh.msh, vortex0.5h.msh,
    vortex0.25h.msh, system/{controlDict,dgSolution,dgSchemes}, constant/transportProperties, 0/{U,p};
  * doubleMach (the limited Euler solver): doubleMach.msh, system/{controlDict,dgSolution,dgSchemes},
    constant/transportProperties, 0/{rho,rhoU,Ener}.
Run only where the reference is present; the copies are committed under tests/golden/cases/.
"""
import os
import shutil
import sys

SRC = "/root/reference/HopeFOAM-0.1/tutorials/DG/2D"
HERE = os.path.dirname(os.path.abspath(__file__))

FILES = {
    "isentropicVortex": ["vortex0256.msh", "vortex1024.msh", "system/controlDict", "system/dgSolution", "system/dgSchemes",
                         "0/rho", "0/rhoU", "0/Ener"],
    "cylinder": ["constant/polyMesh/points", "constant/polyMesh/faces", "constant/polyMesh/owner",
                 "constant/polyMesh/neighbour", "constant/polyMesh/boundary", "system/controlDict",
                 "system/dgSolution", "0/U", "0/p"],
    "pearsonVortex": ["vortexh.msh", "vortex0.5h.msh", "vortex0.25h.msh", "system/controlDict", "system/dgSolution",
                      "system/dgSchemes", "constant/transportProperties", "0/U", "0/p"],
    "doubleMach": ["doubleMach.msh", "system/controlDict", "system/dgSolution", "system/dgSchemes",
                   "constant/transportProperties", "0/rho", "0/rhoU", "0/Ener"],
}


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    for case, files in FILES.items():
        for rel in files:
            dst = os.path.join(HERE, "cases", case, rel)
            os.makedirs(os.path.dirname(dst), exist_ok=True)
            shutil.copyfile(os.path.join(src, case, rel), dst)
            os.chmod(dst, 0o644)
            print(case, rel, os.path.getsize(dst))


if __name__ == "__main__":
    main()
