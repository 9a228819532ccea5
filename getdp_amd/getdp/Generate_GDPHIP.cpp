// The device Generate[A] hook for GetDP.  Generate_System
// (src/kernel/SolvingOperations.cpp:276-500) calls Treatment_Formulation per
// formulation, which walks every element and calls
// Cal_GalerkinTermOfFemEquation per (element, term).  With this file compiled
// in, the loop of SolvingOperations.cpp:442-451 becomes
//
//   if(!(Nbr_Formulation == 1 &&
//        GDP_GenerateFormulation(Formulation_P, DofData_P)))
//     Treatment_Formulation(Formulation_P);
//
// GDP_GenerateFormulation lowers the formulation once -- the GeoData mesh,
// each term's EquationTerm / DefineQuantity / FunctionSpace / Integration /
// Jacobian data and its per-region coefficient, and the DofData's fixed DOFs
// -- into the descriptors of gdp_hip.h, checks that the library numbers the
// DOFs as GetDP's own PRE pass did, and then replaces every later element
// loop by one gdp_fe_generate.  A term outside the device path's forms (the
// classification of Cal_InitGalerkinTermOfFemEquation,
// Cal_GalerkinTermOfFemEquation.cpp:39-316, restricted to GAUSS integration,
// Jacobian Vol/Sur without parameters, order-1 BF_Node / BF_Edge /
// BF_PerpendicularEdge, a per-region constant coefficient, static terms)
// makes it return 0, and GetDP keeps its own element loop.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "LinAlg.h"  // with the HAVE_GDP_HIP branch (LinAlg_GDPHIP.h)

#include "BF.h"
#include "Cal_Quantity.h"
#include "Cal_Value.h"
#include "DofData.h"
#include "F.h"
#include "GeoData.h"
#include "ListUtils.h"
#include "Message.h"
#include "ProData.h"

extern struct Problem Problem_S;
extern struct CurrentData Current;

#if defined(HAVE_GDP_HIP)

int *GDP_CurrentDummyDof(void) { return Current.DofData ? Current.DofData->DummyDof : NULL; }

namespace {

// ---- what one lowered formulation keeps ------------------------------------------

struct Lowered {
  bool supported = false;
  gdp_fe *fe = NULL;
  std::vector<double> xyz;
  std::vector<int> etype, region, conn, edges;
  int nedges = 0;
  std::map<int, int> node_index;  // GetDP node number -> 1-based index (same order)
  std::vector<int> space_dq;      // DefineQuantity index of each space
  std::vector<std::vector<int>> support, fixent, tregions;
  std::vector<std::vector<double>> fixval, tcoef;
  std::vector<gdp_space_desc> spaces;
  std::vector<gdp_term_desc> terms;
};

std::map<const struct Formulation *, Lowered> g_lowered;

// ---- small helpers ------------------------------------------------------------------

int gmsh_type(int t)
{
  if(t == LINE) return GDP_LINE;
  if(t == TRIANGLE) return GDP_TRI;
  if(t == TETRAHEDRON) return GDP_TET;
  return 0;
}

// the region numbers of a REGIONLIST group, or false
bool group_regions(int index, std::vector<int> &out)
{
  if(index < 0) return false;
  struct Group *g = (struct Group *)List_Pointer(Problem_S.Group, index);
  if(g->Type != REGIONLIST || !g->InitialList) return false;
  out.clear();
  for(int k = 0; k < List_Nbr(g->InitialList); k++) {
    int r;
    List_Read(g->InitialList, k, &r);
    out.push_back(r);
  }
  return true;
}

int bf_kind(struct BasisFunction *bf)
{
  if(bf->Function == (void (*)())BF_Node) return GDP_BF_NODE;
  if(bf->Function == (void (*)())BF_Edge) return GDP_BF_EDGE;
  if(bf->Function == (void (*)())BF_PerpendicularEdge) return GDP_BF_PERP_EDGE;
  return -1;
}

// An expression whose value depends on the region only: constants, piecewise
// functions of such, and whole quantities built from constants, other such
// expressions, +, -, *, / and Vector/Tensor constructors.
bool region_constant(int index, int depth = 0);

bool wq_region_constant(List_T *wq, int depth)
{
  if(!wq || depth > 16) return false;
  for(int k = 0; k < List_Nbr(wq); k++) {
    struct WholeQuantity *q = (struct WholeQuantity *)List_Pointer(wq, k);
    switch(q->Type) {
    case WQ_CONSTANT: break;
    case WQ_BINARYOPERATOR:
      if(q->Case.Operator.TypeOperator != OP_PLUS && q->Case.Operator.TypeOperator != OP_MINUS &&
         q->Case.Operator.TypeOperator != OP_TIME && q->Case.Operator.TypeOperator != OP_DIVIDE)
        return false;
      break;
    case WQ_UNARYOPERATOR:
      if(q->Case.Operator.TypeOperator != OP_NEG) return false;
      break;
    case WQ_EXPRESSION:
      if(q->Case.Expression.NbrArguments || !region_constant(q->Case.Expression.Index, depth + 1)) return false;
      break;
    case WQ_BUILTINFUNCTION:
      if(q->Case.Function.Fct != (void (*)())F_Vector && q->Case.Function.Fct != (void (*)())F_Tensor &&
         q->Case.Function.Fct != (void (*)())F_TensorSym && q->Case.Function.Fct != (void (*)())F_TensorDiag)
        return false;
      break;
    default: return false;
    }
  }
  return true;
}

bool region_constant(int index, int depth)
{
  if(index < 0 || depth > 16) return false;
  struct Expression *e = (struct Expression *)List_Pointer(Problem_S.Expression, index);
  switch(e->Type) {
  case CONSTANT: return true;
  case WHOLEQUANTITY: return wq_region_constant(e->Case.WholeQuantity, depth);
  case PIECEWISEFUNCTION: {
    List_T *l = e->Case.PieceWiseFunction.ExpressionPerRegion;
    for(int k = 0; k < List_Nbr(l); k++) {
      struct ExpressionPerRegion *p = (struct ExpressionPerRegion *)List_Pointer(l, k);
      if(!region_constant(p->ExpressionIndex, depth + 1)) return false;
    }
    const int d = e->Case.PieceWiseFunction.ExpressionIndex_Default;
    return d < 0 || region_constant(d, depth + 1);
  }
  default: return false;
  }
}

// a value as the 1, 3 or 9 numbers of gdp_term_desc (tensors full 3x3)
int value_numbers(const struct Value &v, double out[9])
{
  switch(v.Type) {
  case SCALAR: out[0] = v.Val[0]; return 1;
  case VECTOR:
    for(int c = 0; c < 3; c++) out[c] = v.Val[c];
    return 3;
  case TENSOR:
    for(int c = 0; c < 9; c++) out[c] = v.Val[c];
    return 9;
  case TENSOR_SYM: {
    const double *s = v.Val;  // |0 1 2| |s 3 4| |s s 5|
    const double t[9] = {s[0], s[1], s[2], s[1], s[3], s[4], s[2], s[4], s[5]};
    for(int c = 0; c < 9; c++) out[c] = t[c];
    return 9;
  }
  case TENSOR_DIAG: {
    const double *s = v.Val;
    const double t[9] = {s[0], 0., 0., 0., s[1], 0., 0., 0., s[2]};
    for(int c = 0; c < 9; c++) out[c] = t[c];
    return 9;
  }
  default: return 0;
  }
}

// A whole quantity of constants and region-constant expressions, evaluated in
// region r with the reference's own value operations (the stack machine of
// Cal_WholeQuantity, Cal_Quantity.cpp:331-700, for these entry types).
bool eval_wq(List_T *wq, int r, struct Value &out)
{
  std::vector<struct Value> st;
  Current.Region = r;
  for(int k = 0; k < List_Nbr(wq); k++) {
    struct WholeQuantity *q = (struct WholeQuantity *)List_Pointer(wq, k);
    struct Value v;
    memset(&v, 0, sizeof(v));
    switch(q->Type) {
    case WQ_CONSTANT:
      v.Type = SCALAR;
      v.Val[0] = q->Case.Constant;
      st.push_back(v);
      break;
    case WQ_EXPRESSION:
      Get_ValueOfExpressionByIndex(q->Case.Expression.Index, NULL, 0., 0., 0., &v);
      st.push_back(v);
      break;
    case WQ_UNARYOPERATOR:
      if(st.empty()) return false;
      ((void (*)(struct Value *))q->Case.Operator.Function)(&st.back());
      break;
    case WQ_BINARYOPERATOR: {
      if(st.size() < 2) return false;
      struct Value b = st.back();
      st.pop_back();
      ((void (*)(struct Value *, struct Value *, struct Value *))q->Case.Operator.Function)(&st.back(), &b,
                                                                                           &st.back());
      break;
    }
    case WQ_BUILTINFUNCTION: {
      const int na = q->Case.Function.NbrArguments;
      if((int)st.size() < na) return false;
      std::vector<struct Value> args(st.end() - na, st.end());
      st.resize(st.size() - na);
      ((void (*)(struct Function *, struct Value *, struct Value *))q->Case.Function.Fct)(&q->Case.Function,
                                                                                         args.data(), &v);
      st.push_back(v);
      break;
    }
    default: return false;
    }
  }
  if(st.size() != 1) return false;
  out = st[0];
  return true;
}

// Gauss points of an integration method for an element type (GetDP type), or 0
int gauss_points(int method, int getdp_type)
{
  if(method < 0) return 0;
  struct IntegrationMethod *m = (struct IntegrationMethod *)List_Pointer(Problem_S.IntegrationMethod, method);
  if(m->CriterionIndex >= 0 || List_Nbr(m->IntegrationCase) != 1) return 0;
  struct IntegrationCase *c = (struct IntegrationCase *)List_Pointer(m->IntegrationCase, 0);
  if(c->Type != GAUSS || c->SubType != STANDARD) return 0;
  for(int k = 0; k < List_Nbr(c->Case); k++) {
    struct Quadrature *q = (struct Quadrature *)List_Pointer(c->Case, k);
    if(q->ElementType == getdp_type) return q->NumberOfPoints;
  }
  return 0;
}

// -1: unsupported; 0: Jacobian Vol; 1: Jacobian Sur
int jacobian_kind(int method)
{
  if(method < 0) return -1;
  struct JacobianMethod *m = (struct JacobianMethod *)List_Pointer(Problem_S.JacobianMethod, method);
  int kind = -2;
  for(int k = 0; k < List_Nbr(m->JacobianCase); k++) {
    struct JacobianCase *c = (struct JacobianCase *)List_Pointer(m->JacobianCase, k);
    if(c->NbrParameters || c->CoefficientIndex >= 0) return -1;
    const int j = c->TypeJacobian == JACOBIAN_VOL ? 0 : c->TypeJacobian == JACOBIAN_SUR ? 1 : -1;
    if(j < 0 || (kind != -2 && kind != j)) return -1;
    kind = j;
  }
  return kind == -2 ? -1 : kind;
}

// ---- lowering ------------------------------------------------------------------------

bool lower_mesh(Lowered &L)
{
  const int nn = Geo_GetNbrGeoNodes(), ne = Geo_GetNbrGeoElements();
  std::vector<std::pair<int, const struct Geo_Node *>> nodes(nn);
  for(int i = 0; i < nn; i++) {
    const struct Geo_Node *p = Geo_GetGeoNode(i);
    nodes[i] = {p->Num, p};
  }
  std::sort(nodes.begin(), nodes.end(),
            [](const std::pair<int, const struct Geo_Node *> &a, const std::pair<int, const struct Geo_Node *> &b) {
              return a.first < b.first;
            });
  L.xyz.resize(3 * (size_t)nn);
  for(int i = 0; i < nn; i++) {
    L.node_index[nodes[i].first] = i + 1;
    L.xyz[3 * i] = nodes[i].second->x;
    L.xyz[3 * i + 1] = nodes[i].second->y;
    L.xyz[3 * i + 2] = nodes[i].second->z;
  }
  // elements: GeoData keeps them sorted by number (GeoData.cpp:1018)
  L.etype.resize(ne);
  L.region.resize(ne);
  L.conn.assign(4 * (size_t)ne, 0);
  bool with_edges = false;
  for(int e = 0; e < ne; e++) with_edges |= Geo_GetGeoElement(e)->NbrEdges > 0;
  if(with_edges) L.edges.assign(6 * (size_t)ne, 0);
  for(int e = 0; e < ne; e++) {
    struct Geo_Element *g = Geo_GetGeoElement(e);
    const int t = gmsh_type(g->Type);
    if(!t || g->NbrNodes > 4 || g->Region < 0) return false;
    L.etype[e] = t;
    L.region[e] = g->Region;
    for(int k = 0; k < g->NbrNodes; k++) L.conn[4 * (size_t)e + k] = L.node_index[g->NumNodes[k]];
    if(with_edges)
      for(int k = 0; k < g->NbrEdges && k < 6; k++) {
        L.edges[6 * (size_t)e + k] = g->NumEdges[k];
        L.nedges = std::max(L.nedges, abs(g->NumEdges[k]));
      }
  }
  return true;
}

// the space of DefineQuantity dq (added on first use); -1 if unsupported
int space_of(Lowered &L, struct Formulation *F, int dq, struct DofData *DD)
{
  for(size_t s = 0; s < L.space_dq.size(); s++)
    if(L.space_dq[s] == dq) return (int)s;
  struct DefineQuantity *q = (struct DefineQuantity *)List_Pointer(F->DefineQuantity, dq);
  if(q->Type != LOCALQUANTITY || List_Nbr(q->IndexInFunctionSpace)) return -1;
  struct FunctionSpace *fs = (struct FunctionSpace *)List_Pointer(Problem_S.FunctionSpace, q->FunctionSpaceIndex);
  if(List_Nbr(fs->BasisFunction) != 1 || List_Nbr(fs->GlobalQuantity)) return -1;
  struct BasisFunction *bf = (struct BasisFunction *)List_Pointer(fs->BasisFunction, 0);
  const int kind = bf_kind(bf);
  std::vector<int> sup;
  if(kind < 0 || !group_regions(bf->SupportIndex, sup)) return -1;
  // fixed DOFs of this space as GetDP's PRE pass left them (the first
  // constraint holding an entity won there, DofData.cpp:1137-1151)
  std::vector<int> fe_;
  std::vector<double> fv;
  if(!DD->DofList) return -1;
  for(int k = 0; k < List_Nbr(DD->DofList); k++) {
    struct Dof *d = (struct Dof *)List_Pointer(DD->DofList, k);
    if(d->NumType != bf->Num) continue;
    if(d->Type == DOF_FIXED) {
      fe_.push_back(kind == GDP_BF_EDGE ? d->Entity : L.node_index[d->Entity]);
      fv.push_back(d->Val.s);
    }
    else if(d->Type != DOF_UNKNOWN)
      return -1;  // links, associations, initial values: GetDP's own loop
  }
  L.space_dq.push_back(dq);
  L.support.push_back(sup);
  L.fixent.push_back(fe_);
  L.fixval.push_back(fv);
  gdp_space_desc sd;
  sd.bf = kind;
  sd.numtype = bf->Num;
  L.spaces.push_back(sd);
  return (int)L.spaces.size() - 1;
}

bool lower_terms(Lowered &L, struct Formulation *F, struct DofData *DD)
{
  for(int k = 0; k < List_Nbr(F->Equation); k++) {
    struct EquationTerm *et = (struct EquationTerm *)List_Pointer(F->Equation, k);
    if(et->Type != GALERKIN) return false;
    const auto &lt = et->Case.LocalTerm;
    if(lt.Term.TypeTimeDerivative != NODT_ || lt.SubRegion >= 0 || lt.Full_Matrix ||
       lt.ExpressionIndexForMetricTensor >= 0 || lt.MatrixIndex >= 0)
      return false;
    const int se = space_of(L, F, lt.Term.DefineQuantityIndexEqu, DD);
    if(se < 0) return false;
    const int bf = L.spaces[se].bf;
    const bool has_dof = lt.Term.DefineQuantityIndexDof >= 0;
    int sd = -1;
    if(has_dof) {
      sd = space_of(L, F, lt.Term.DefineQuantityIndexDof, DD);
      if(sd < 0 || L.spaces[sd].bf != bf || lt.Term.TypeOperatorDof != lt.Term.TypeOperatorEqu) return false;
      if(lt.Term.CanonicalWholeQuantity != CWQ_DOF && lt.Term.CanonicalWholeQuantity != CWQ_EXP_TIME_DOF)
        return false;
    }
    const int op = lt.Term.TypeOperatorEqu;
    const bool deriv = op == EXTDER || op == GRAD || op == CURL;
    if(!deriv && op != NOOP) return false;
    int kind;
    if(has_dof) kind = deriv ? (bf == GDP_BF_NODE ? GDP_T_GRAD_GRAD : GDP_T_CURL_CURL)
                             : (bf == GDP_BF_NODE ? GDP_T_NODE_MASS : GDP_T_EDGE_MASS);
    else if(!deriv) kind = bf == GDP_BF_NODE ? GDP_T_SRC_NODE : GDP_T_SRC_EDGE;
    else return false;
    std::vector<int> regs;
    if(!group_regions(lt.InIndex, regs)) return false;
    const int jac = jacobian_kind(lt.JacobianMethodIndex);
    if(jac < 0) return false;
    // Gauss points: the same count for every element type the term meets
    int npts = 0;
    for(size_t e = 0; e < L.etype.size(); e++) {
      if(std::find(regs.begin(), regs.end(), L.region[e]) == regs.end()) continue;
      const int gt = L.etype[e] == GDP_TRI ? TRIANGLE : L.etype[e] == GDP_TET ? TETRAHEDRON : LINE;
      const int p = gauss_points(lt.IntegrationMethodIndex, gt);
      if(!p || (npts && p != npts)) return false;
      npts = p;
    }
    if(!npts) return false;
    // the per-region coefficient, evaluated once per region
    const int maxreg = *std::max_element(L.region.begin(), L.region.end());
    std::vector<double> coef;
    int ncoef = 0;
    for(int r : regs) {
      struct Value v;
      memset(&v, 0, sizeof(v));
      if(has_dof) {
        if(lt.Term.CanonicalWholeQuantity == CWQ_DOF) {
          v.Type = SCALAR;
          v.Val[0] = 1.;
        }
        else {
          if(!region_constant(lt.Term.ExpressionIndexForCanonical)) return false;
          Current.Region = r;
          Get_ValueOfExpressionByIndex(lt.Term.ExpressionIndexForCanonical, NULL, 0., 0., 0., &v);
        }
      }
      else if(!wq_region_constant(lt.Term.WholeQuantity, 0) || !eval_wq(lt.Term.WholeQuantity, r, v))
        return false;
      double num[9];
      const int nc = value_numbers(v, num);
      if(!nc || (ncoef && nc != ncoef)) return false;
      if(!ncoef) {
        ncoef = nc;
        coef.assign((size_t)(maxreg + 1) * ncoef, 0.);
      }
      if(r < 0 || r > maxreg) continue;
      for(int c = 0; c < ncoef; c++) coef[(size_t)r * ncoef + c] = num[c];
    }
    L.tregions.push_back(regs);
    L.tcoef.push_back(coef);
    gdp_term_desc t;
    t.kind = kind;
    t.equ = se;
    t.dof = sd;
    t.npoints = npts;
    t.jac_sur = jac;
    t.ncoef = ncoef;
    t.coef_nreg = maxreg + 1;
    L.terms.push_back(t);
  }
  return !L.terms.empty();
}

// the library's numbering must be GetDP's own (DofData's unknowns)
bool same_numbering(Lowered &L, struct DofData *DD)
{
  int N = 0;
  long nnz = 0;
  if(gdp_fe_num_dofs(L.fe, &N, &nnz) || N != DD->NbrDof) return false;
  const int nmax = std::max((int)L.node_index.size(), L.nedges);
  std::vector<std::vector<int>> tab(L.spaces.size(), std::vector<int>((size_t)nmax + 1));
  for(size_t s = 0; s < L.spaces.size(); s++)
    if(gdp_fe_dof_table(L.fe, (int)s, tab[s].data(), nmax)) return false;
  for(int k = 0; k < List_Nbr(DD->DofList); k++) {
    struct Dof *d = (struct Dof *)List_Pointer(DD->DofList, k);
    if(d->Type != DOF_UNKNOWN) continue;
    for(size_t s = 0; s < L.spaces.size(); s++) {
      if(L.spaces[s].numtype != d->NumType) continue;
      const int ent = L.spaces[s].bf == GDP_BF_EDGE ? d->Entity : L.node_index[d->Entity];
      if(ent < 1 || ent > nmax || tab[s][ent] != d->Case.Unknown.NumDof) return false;
    }
  }
  return true;
}

bool lower(Lowered &L, struct Formulation *F, struct DofData *DD)
{
  if(!lower_mesh(L) || !lower_terms(L, F, DD)) return false;
  for(size_t s = 0; s < L.spaces.size(); s++) {
    L.spaces[s].nsupport = (int)L.support[s].size();
    L.spaces[s].support = L.support[s].data();
    L.spaces[s].nfixed = (int)L.fixent[s].size();
    L.spaces[s].fixed_entity = L.fixent[s].data();
    L.spaces[s].fixed_value = L.fixval[s].data();
  }
  for(size_t t = 0; t < L.terms.size(); t++) {
    L.terms[t].nregions = (int)L.tregions[t].size();
    L.terms[t].regions = L.tregions[t].data();
    L.terms[t].coef = L.tcoef[t].data();
  }
  gdp_mesh_desc m;
  m.nnodes = (int)L.node_index.size();
  m.xyz = L.xyz.data();
  m.nelem = (int)L.etype.size();
  m.etype = L.etype.data();
  m.region = L.region.data();
  m.conn = L.conn.data();
  m.edges = L.edges.empty() ? NULL : L.edges.data();
  m.nedges = L.edges.empty() ? 0 : L.nedges;
  const int rc = gdp_fe_create(&m, (int)L.spaces.size(), L.spaces.data(), (int)L.terms.size(), L.terms.data(),
                               &L.fe);
  if(rc == GDP_E_UNSUPPORTED) return false;
  if(rc) {
    Message::Warning("Device assembly unavailable: %s", gdp_last_error());
    return false;
  }
  if(!same_numbering(L, DD)) {
    Message::Warning("Device assembly: DOF numbering differs from the PRE pass; using the element loop");
    gdp_fe_destroy(L.fe);
    L.fe = NULL;
    return false;
  }
  return true;
}

}  // namespace

// 1 when Generate[A] of this formulation ran on the device (A and b of the
// DofData zeroed and assembled), 0 when GetDP's element loop must do it.
int GDP_GenerateFormulation(struct Formulation *Formulation_P, struct DofData *DofData_P)
{
  if(!Formulation_P || !DofData_P || Formulation_P->Type != FEMEQUATION) return 0;
  if(DofData_P->NbrHar != 1 || DofData_P->Flag_Only || DofData_P->Flag_RHS) return 0;
  if(!DofData_P->Flag_Init[0] || DofData_P->Flag_Init[0] > 1) return 0;  // A and b only
  auto it = g_lowered.find(Formulation_P);
  if(it == g_lowered.end()) {
    Lowered &L = g_lowered[Formulation_P];
    L.supported = lower(L, Formulation_P, DofData_P);
    it = g_lowered.find(Formulation_P);
  }
  if(!it->second.supported) return 0;
  const int rc = gdp_fe_generate(it->second.fe, DofData_P->A.M, DofData_P->b.V);
  if(rc) {
    Message::Error("Generate on the device: %s", gdp_last_error());
    return 0;
  }
  DofData_P->b.mirror_state = 0;
  return 1;
}

#endif
