"""MULTIGRIDParams XML (quda_multigrid_params.cc:44-235) read into cmx_mg_params, host only."""
import pytest

from chroma_amd import lib as cmx


def test_mg_params_from_reference_xml():
    xml = """<InvertParam><MULTIGRIDParams>
      <Verbosity>true</Verbosity><Precision>SINGLE</Precision><Reconstruct>RECONS_NONE</Reconstruct>
      <NullVectors>24</NullVectors><GenerateNullspace>true</GenerateNullspace>
      <Pre-SmootherApplications>0</Pre-SmootherApplications><Post-SmootherApplications>6</Post-SmootherApplications>
      <CoarseResidual>0.1 0.25</CoarseResidual><MaxCoarseIterations>10 8</MaxCoarseIterations>
      <RelaxationOmegaMG>0.85</RelaxationOmegaMG><PrecondGCRNKrylov>16</PrecondGCRNKrylov>
      <Blocking><elem>4 4 4 2</elem></Blocking><CycleType>MG_VCYCLE</CycleType>
    </MULTIGRIDParams></InvertParam>"""
    p = cmx.mg_params_from_xml(xml)
    assert list(p.blocking) == [4, 4, 4, 2]
    assert (p.nvec, p.nu_pre, p.nu_post, p.coarse_nkrylov, p.verbose) == (24, 0, 6, 16, 1)
    assert p.omega == pytest.approx(0.85) and p.coarse_tol == pytest.approx(0.25) and p.coarse_maxit == 8
    assert p.prec == cmx.SINGLE
    d = cmx.mg_params()
    assert list(d.blocking) == [4, 4, 4, 4] and d.nvec == 16 and d.coarse_maxit == 12 and d.setup_maxit == 500
    assert d.prec == cmx.SINGLE
    for name, want in (("HALF", cmx.HALF), ("DOUBLE", cmx.DOUBLE), ("DEFAULT", cmx.SINGLE)):
        q = cmx.mg_params_from_xml(f"<MULTIGRIDParams><Precision>{name}</Precision>"
                                   "<Blocking><elem>4 4 4 4</elem></Blocking></MULTIGRIDParams>")
        assert q.prec == want
    with pytest.raises(ValueError):
        cmx.mg_params_from_xml("<MULTIGRIDParams><Precision>QUARTER</Precision>"
                               "<Blocking><elem>4 4 4 4</elem></Blocking></MULTIGRIDParams>")
    with pytest.raises(ValueError):
        cmx.mg_params_from_xml("<MULTIGRIDParams><Blocking><elem>4 4 4 4</elem><elem>2 2 2 2</elem>"
                               "</Blocking></MULTIGRIDParams>")
